"""Differentiable LSCM solve of one chart (positions -> uv) and its vector-Jacobian product.

Pins are frozen constants of the positions; the default path (angle_mix = 1) is assumed.  The reverse pass
is one extra PCG solve with the same SPD matrix followed by a device gather through the cotan assembly
(include/ruv_lscm.h: ruv_lscm_forward / ruv_lscm_vjp).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .solver import Precond


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _setup(L):
    if getattr(L, "_ruv_diff_ready", False):
        return L
    P = C.c_void_p
    L.ruv_lscm_forward.restype = C.c_int
    L.ruv_lscm_forward.argtypes = [P, C.c_int64, P, C.c_int64, P, P, C.c_int64, C.c_int32, C.c_int32, C.c_double, P,
                                   P, P]
    L.ruv_lscm_vjp.restype = C.c_int
    L.ruv_lscm_vjp.argtypes = [P, C.c_int64, P, C.c_int64, P, P, C.c_int64, P, P, C.c_double, C.c_int32, C.c_int32,
                               C.c_double, P, P]
    L._ruv_diff_ready = True
    return L


class ChartSolve:
    """One chart's topology and frozen pins; forward / vjp take positions (nv, 3)."""

    def __init__(self, tris, pin_ids, pin_uv, precond=Precond.JACOBI, mixed=False, tol=1e-12):
        self.tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        self.pin_ids = np.ascontiguousarray(pin_ids, np.int32)
        self.pin_uv = np.ascontiguousarray(pin_uv, np.float64).reshape(-1, 2)
        self.precond = int(precond)
        self.mode = 1 if mixed else 0
        self.tol = tol
        self.L = _setup(_lib.lib())

    def forward(self, p3):
        p3 = np.ascontiguousarray(p3, np.float64).reshape(-1, 3)
        uv = np.zeros((len(p3), 2))
        e = C.c_double(0)
        info = _lib.SolveInfo()
        _lib.check(self.L.ruv_lscm_forward(_p(p3), len(p3), _p(self.tris), len(self.tris), _p(self.pin_ids),
                                           _p(self.pin_uv), len(self.pin_ids), self.precond, self.mode, self.tol,
                                           _p(uv), C.byref(e), C.byref(info)))
        return uv, e.value

    def vjp(self, p3, uv, g_uv, energy_weight=0.0):
        p3 = np.ascontiguousarray(p3, np.float64).reshape(-1, 3)
        uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
        g = np.ascontiguousarray(g_uv, np.float64).reshape(-1, 2)
        grad = np.zeros_like(p3)
        info = _lib.SolveInfo()
        _lib.check(self.L.ruv_lscm_vjp(_p(p3), len(p3), _p(self.tris), len(self.tris), _p(self.pin_ids),
                                       _p(self.pin_uv), len(self.pin_ids), _p(uv), _p(g), energy_weight,
                                       self.precond, self.mode, self.tol, _p(grad), C.byref(info)))
        return grad


class BatchSolve:
    """Differentiable solve of many charts at once.  The disjoint union of the charts is one chart for the
    device assembly, the solve and the adjoint gather (its L is block-diagonal by chart), so one forward and
    one adjoint solve serve the whole batch.  charts: iterable of (nv, tris, pin_ids, pin_uv), chart-local;
    positions and gradients are the charts' (nv, 3) blocks stacked in order."""

    def __init__(self, charts, precond=Precond.JACOBI, mixed=False, tol=1e-12):
        tris, pins, puv, off = [], [], [], [0]
        for nv, t, pid, pu in charts:
            base = off[-1]
            tris.append(np.asarray(t, np.int64).reshape(-1, 3) + base)
            pins.append(np.asarray(pid, np.int64) + base)
            puv.append(np.asarray(pu, np.float64).reshape(-1, 2))
            off.append(base + int(nv))
        self.offsets = np.asarray(off, np.int64)
        self.cs = ChartSolve(np.concatenate(tris).astype(np.uint32), np.concatenate(pins).astype(np.int32),
                             np.concatenate(puv), precond=precond, mixed=mixed, tol=tol)

    def forward(self, p3):
        return self.cs.forward(p3)

    def vjp(self, p3, uv, g_uv, energy_weight=0.0):
        return self.cs.vjp(p3, uv, g_uv, energy_weight)

    def split(self, a):
        """Per-chart blocks of a stacked (nv_total, k) array."""
        return [a[self.offsets[c]:self.offsets[c + 1]] for c in range(len(self.offsets) - 1)]


class DiffBatch:
    """The differentiable batch kept on the device (include/ruv_lscm.h: ruv_lscm_diff_*): the charts' disjoint
    union is assembled and solved by the batched engine (each chart its own PCG), topology and structure built
    once; forward(p3) -> (uv, energy); vjp(g_uv, energy_weight) on the state of the last forward."""

    def __init__(self, p3, charts, precond=Precond.AMG, mixed=True, tol=1e-10):
        L = _lib.lib()
        if not getattr(L, "_ruv_diffb_ready", False):
            P = C.c_void_p
            L.ruv_lscm_diff_create.restype = C.c_int
            L.ruv_lscm_diff_create.argtypes = [P, C.c_int64, P, C.c_int64, P, P, C.c_int64, P, C.c_int32, C.c_int32,
                                               C.c_int32, C.c_double, P]
            L.ruv_lscm_diff_forward.restype = C.c_int
            L.ruv_lscm_diff_forward.argtypes = [P, P, P, P, P]
            L.ruv_lscm_diff_vjp.restype = C.c_int
            L.ruv_lscm_diff_vjp.argtypes = [P, P, C.c_double, P, P]
            L.ruv_lscm_diff_destroy.restype = None
            L.ruv_lscm_diff_destroy.argtypes = [P]
            L._ruv_diffb_ready = True
        self.L = L
        tris, pins, puv, off = [], [], [], [0]
        for nv, t, pid, pu in charts:
            base = off[-1]
            tris.append(np.asarray(t, np.int64).reshape(-1, 3) + base)
            pins.append(np.asarray(pid, np.int64) + base)
            puv.append(np.asarray(pu, np.float64).reshape(-1, 2))
            off.append(base + int(nv))
        self.offsets = np.asarray(off, np.int64)
        self.nv = int(off[-1])
        T = np.ascontiguousarray(np.concatenate(tris), np.uint32)
        pid = np.ascontiguousarray(np.concatenate(pins), np.int32)
        pu = np.ascontiguousarray(np.concatenate(puv), np.float64)
        p3 = np.ascontiguousarray(p3, np.float64).reshape(-1, 3)
        self.h = C.c_void_p()
        _lib.check(L.ruv_lscm_diff_create(_p(p3), self.nv, _p(T), len(T), _p(pid), _p(pu), len(pid), _p(self.offsets),
                                          len(self.offsets) - 1, int(precond), 1 if mixed else 0, tol,
                                          C.byref(self.h)))

    def forward(self, p3):
        p3 = np.ascontiguousarray(p3, np.float64).reshape(-1, 3)
        uv = np.zeros((self.nv, 2))
        e = C.c_double(0)
        info = _lib.SolveInfo()
        _lib.check(self.L.ruv_lscm_diff_forward(self.h, _p(p3), _p(uv), C.byref(e), C.byref(info)))
        self.last_info = info
        return uv, e.value

    def vjp(self, p3, uv, g_uv, energy_weight=0.0):
        """p3 and uv must be those of the last forward (kept on the device)."""
        g = np.ascontiguousarray(g_uv, np.float64).reshape(-1, 2)
        grad = np.zeros((self.nv, 3))
        info = _lib.SolveInfo()
        _lib.check(self.L.ruv_lscm_diff_vjp(self.h, _p(g), energy_weight, _p(grad), C.byref(info)))
        return grad

    def split(self, a):
        return [a[self.offsets[c]:self.offsets[c + 1]] for c in range(len(self.offsets) - 1)]

    def close(self):
        if self.h:
            self.L.ruv_lscm_diff_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
