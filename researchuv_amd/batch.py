"""Batched LSCM solves: many independent charts in shared launches (each chart its own PCG).

`LscmBatchBuilder` runs the host side (weld, segment, assemble) over many meshes on worker threads;
`Batch` owns the device-resident block-diagonal system and solves every chart to its own tolerance.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Optional, Tuple

import numpy as np

from . import _lib
from .solver import Precond


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _setup(L):
    if getattr(L, "_ruv_batch_ready", False):
        return L
    P = C.c_void_p
    L.ruv_batch_create.restype = C.c_int
    L.ruv_batch_create.argtypes = [C.c_int64, P, P, P, C.c_int32, P, C.c_int32, C.c_int32, C.c_int32, P]
    L.ruv_batch_solve_dev.restype = C.c_int
    L.ruv_batch_solve_dev.argtypes = [P, P, P, C.c_uint64, C.c_double, P, P, P]
    L.ruv_batch_solve.restype = C.c_int
    L.ruv_batch_solve.argtypes = [P, P, P, C.c_uint64, C.c_double, P, P, P]
    L.ruv_batch_spmv_time.restype = C.c_int
    L.ruv_batch_spmv_time.argtypes = [P, C.c_int32, P, P, P]
    L.ruv_batch_schedule.restype = C.c_int
    L.ruv_batch_kernel_timing.restype = C.c_int
    L.ruv_batch_precond_info.restype = C.c_int
    L.ruv_batch_setup_precond.restype = C.c_int
    L.ruv_batch_precond_levels.restype = C.c_int
    L.ruv_batch_precond_levels.argtypes = [P, P, C.c_int32]
    L.ruv_batch_precond_fallbacks.restype = C.c_int
    L.ruv_batch_precond_fallbacks.argtypes = [P, P, P, P]
    L.ruv_batch_coarsest_block.restype = C.c_int
    L.ruv_batch_coarsest_block.argtypes = [P, C.c_int32, P, P, P]
    L.ruv_batch_amg_level.restype = C.c_int
    L.ruv_batch_amg_level.argtypes = [P, C.c_int32, P, P, P, P, P, P]
    L.ruv_batch_setup_precond.argtypes = [P]
    L.ruv_batch_precond_info.argtypes = [P, P, P, P]
    L.ruv_batch_kernel_timing.argtypes = [P, C.c_int32]
    L.ruv_batch_kernel_stats.restype = C.c_int
    L.ruv_batch_kernel_stats.argtypes = [P, P, P, P, P]
    L.ruv_batch_step_profile.restype = C.c_int
    L.ruv_batch_step_profile.argtypes = [P, C.c_int32, C.c_int32, P, P, P, P, P, P]
    L.ruv_batch_schedule.argtypes = [P, P, P, P, P, P, P, P]
    L.ruv_batch_band_info.restype = C.c_int
    L.ruv_batch_band_info.argtypes = [P, P, P]
    L.ruv_batch_destroy.restype = None
    L.ruv_batch_destroy.argtypes = [P]
    L.ruv_lscm_batch_build.restype = C.c_int
    L.ruv_lscm_batch_build.argtypes = [C.c_int32, P, P, P, P, C.c_double, C.c_double, C.c_double, C.c_int32, P]
    L.ruv_lscm_batch_sizes.restype = C.c_int
    L.ruv_lscm_batch_sizes.argtypes = [P, P]
    L.ruv_lscm_batch_arrays.restype = C.c_int
    L.ruv_lscm_batch_arrays.argtypes = [P, P, P, P, P, P, P, P, P]
    L.ruv_batch_create_lscm.restype = C.c_int
    L.ruv_batch_create_lscm.argtypes = [P, C.c_int32, C.c_int32, P]
    L.ruv_lscm_batch_destroy.restype = None
    L.ruv_lscm_batch_destroy.argtypes = [P]
    L._ruv_batch_ready = True
    return L


def lib():
    return _setup(_lib.lib())


class Batch:
    """Device-resident block-diagonal batch of L_c (x) I_R systems."""

    def __init__(self, h, nrows: int, nchart: int, R: int = 2):
        self.h = h
        self.nrows = nrows
        self.nchart = nchart
        self.R = R

    @classmethod
    def from_csr(cls, row_ptr, cols, vals, chart_row_off, precond=Precond.JACOBI, mixed=True, R=2):
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        cols = np.ascontiguousarray(cols, np.int32)
        vals = np.ascontiguousarray(vals, np.float64)
        off = np.ascontiguousarray(chart_row_off, np.int64)
        h = C.c_void_p()
        _lib.check(lib().ruv_batch_create(len(row_ptr) - 1, _p(row_ptr), _p(cols), _p(vals), len(off) - 1, _p(off), R,
                                          int(precond), 1 if mixed else 0, C.byref(h)))
        return cls(h, len(row_ptr) - 1, len(off) - 1, R)

    @classmethod
    def from_charts(cls, systems, precond=Precond.JACOBI, mixed=True):
        """systems: dicts with l_vals/l_cols/l_ptr/rhs (Segmentation.system).  Returns (batch, rhs)."""
        ptrs, cols, vals, rhs, off = [np.zeros(1, np.int64)], [], [], [], [0]
        nnz = 0
        for s in systems:
            base = off[-1]
            ptrs.append(s["l_ptr"][1:].astype(np.int64) + nnz)
            cols.append(s["l_cols"].astype(np.int64) + base)
            vals.append(s["l_vals"])
            rhs.append(s["rhs"])
            nnz += len(s["l_vals"])
            off.append(base + len(s["l_ptr"]) - 1)
        b = cls.from_csr(np.concatenate(ptrs), np.concatenate(cols).astype(np.int32), np.concatenate(vals), off,
                         precond, mixed)
        return b, np.concatenate(rhs)

    def solve(self, b, max_iter=16384, tol=1e-8):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        info = _lib.SolveInfo()
        it = np.zeros(self.nchart, np.uint64)
        res = np.zeros(self.nchart)
        _lib.check(lib().ruv_batch_solve(self.h, _p(b), _p(x), max_iter, tol, C.byref(info), _p(it), _p(res)))
        return x, info, it, res

    def solve_dev(self, b_ptr: int, x_ptr: int, max_iter=16384, tol=1e-8, per_chart=False):
        info = _lib.SolveInfo()
        it = np.zeros(self.nchart, np.uint64) if per_chart else None
        res = np.zeros(self.nchart) if per_chart else None
        _lib.check(lib().ruv_batch_solve_dev(self.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), max_iter, tol,
                                             C.byref(info), _p(it), _p(res)))
        return (info, it, res) if per_chart else info

    def spmv_time(self, reps=50, stored=False):
        """Mean seconds per launch of the SpMV kernel over the whole batch and its algorithmic bytes (nonzeros
        only); with stored=True also the bytes as stored (padding slots included)."""
        sec = C.c_double(0)
        byt = C.c_double(0)
        sto = C.c_double(0)
        _lib.check(lib().ruv_batch_spmv_time(self.h, reps, C.byref(sec), C.byref(byt), C.byref(sto)))
        return (sec.value, byt.value, sto.value) if stored else (sec.value, byt.value)

    def schedule(self) -> dict:
        """The autotuned schedule: rows per workgroup tile, bytes per column index, stored ELL entries."""
        rows = C.c_int32(0)
        cb = C.c_int32(0)
        ne = C.c_int64(0)
        nc = C.c_int64(0)
        st = C.c_int32(0)
        ns = C.c_int32(0)
        nt = C.c_int32(0)
        _lib.check(lib().ruv_batch_schedule(self.h, C.byref(rows), C.byref(cb), C.byref(ne), C.byref(nc),
                                            C.byref(st), C.byref(ns), C.byref(nt)))
        return dict(rows_per_tile=rows.value, col_bytes=cb.value, ell_entries=ne.value, col_entries=nc.value,
                    stencil_slices=st.value, slices=ns.value, nontemporal=nt.value)

    def band_info(self) -> dict:
        """Charts kept in band form (grid charts with a symmetric operator) and the values that form stores."""
        nb = C.c_int32(0)
        nv = C.c_int64(0)
        _lib.check(lib().ruv_batch_band_info(self.h, C.byref(nb), C.byref(nv)))
        return dict(band_charts=nb.value, band_values=nv.value)

    def setup_precond(self):
        """AMG: rebuild the hierarchy on the device (no-op for the element-wise preconditioners)."""
        _lib.check(lib().ruv_batch_setup_precond(self.h))

    def precond_info(self) -> dict:
        """AMG: levels (finest included), device setup seconds, rows of all coarse levels."""
        lv = C.c_int32(0)
        sec = C.c_double(0)
        rows = C.c_int64(0)
        _lib.check(lib().ruv_batch_precond_info(self.h, C.byref(lv), C.byref(sec), C.byref(rows)))
        per = np.zeros(64, np.int64)
        nl = lib().ruv_batch_precond_levels(self.h, _p(per), 64)
        fb = np.zeros(3, np.int64)
        _lib.check(lib().ruv_batch_precond_fallbacks(self.h, _p(fb[0:1]), _p(fb[1:2]), _p(fb[2:3])))
        return dict(levels=lv.value, setup_seconds=sec.value, coarse_rows=rows.value,
                    level_rows=[int(v) for v in per[:max(nl, 0)]],
                    coarsest_diagonal_charts={"not_spd": int(fb[0]), "coupled": int(fb[1]), "too_large": int(fb[2])})

    def coarsest_block(self, chart: int):
        """AMG: chart's dense block of the coarsest operator and the inverse the device factorisation made."""
        m = C.c_int64(0)
        _lib.check(lib().ruv_batch_coarsest_block(self.h, chart, C.byref(m), None, None))
        blk = np.zeros((m.value, m.value))
        inv = np.zeros((m.value, m.value))
        if m.value:
            _lib.check(lib().ruv_batch_coarsest_block(self.h, chart, C.byref(m), _p(blk), _p(inv)))
        return blk, inv

    def amg_level(self, level: int) -> dict:
        """AMG: level `level` of the device hierarchy in its stored order: op (scipy CSR, None for level 0), agg
        (row -> next-level row; None on the coarsest), chart (row -> chart), rows of the next level."""
        import scipy.sparse as sp

        sz = np.zeros(3, np.int64)
        _lib.check(lib().ruv_batch_amg_level(self.h, level, _p(sz), None, None, None, None, None))
        n, nz, nn = (int(v) for v in sz)
        ptr, cols, vals = np.zeros(n + 1, np.int64), np.zeros(max(nz, 1), np.int32), np.zeros(max(nz, 1))
        agg = np.zeros(n, np.int32) if nn else None
        chart = np.zeros(n, np.int32)
        _lib.check(lib().ruv_batch_amg_level(self.h, level, _p(sz), _p(ptr), _p(cols), _p(vals), _p(agg), _p(chart)))
        op = sp.csr_matrix((vals[:nz], cols[:nz], ptr), shape=(n, n)) if level > 0 else None
        return dict(op=op, agg=agg, chart=chart, n_next=nn)

    def kernel_timing(self, enable=True):
        """Bracket every step-kernel launch of the following solves with HIP events."""
        _lib.check(lib().ruv_batch_kernel_timing(self.h, 1 if enable else 0))

    def kernel_stats(self) -> dict:
        """Launches, summed device seconds, algorithmic bytes (nonzeros only) and bytes as stored (padding
        included) of the SpMV kernel since the last call."""
        n = C.c_int64(0)
        sec = C.c_double(0)
        byt = C.c_double(0)
        sto = C.c_double(0)
        _lib.check(lib().ruv_batch_kernel_stats(self.h, C.byref(n), C.byref(sec), C.byref(byt), C.byref(sto)))
        return dict(launches=n.value, seconds=sec.value, bytes=byt.value, stored_bytes=sto.value)

    STEP_KERNELS = ("kb_dir", "kb_spmv", "kb_update", "kc_resid0", "kc_restrict", "kc_down", "kc_coarsest",
                    "kc_prolong", "kc_post", "kc_prolong0", "kc_post0")

    def step_profile(self, reps=5) -> list:
        """AMG: per kernel launch of the preconditioned step over the whole batch (every chart running), in launch
        order: kernel, level, mean seconds per launch, algorithmic bytes and bytes as stored of one launch.  The
        solver's state is overwritten (the next solve starts afresh)."""
        cap = 256
        kind, level = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        sec, byt, sto = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        cnt = C.c_int32(0)
        _lib.check(lib().ruv_batch_step_profile(self.h, reps, cap, _p(kind), _p(level), _p(sec), _p(byt), _p(sto),
                                                C.byref(cnt)))
        n = min(cnt.value, cap)
        return [dict(kernel=self.STEP_KERNELS[kind[j]], level=int(level[j]), seconds=float(sec[j]),
                     bytes=float(byt[j]), stored_bytes=float(sto[j])) for j in range(n)]

    def close(self):
        if self.h:
            lib().ruv_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LscmBatchBuilder:
    """Host assembly of many meshes (worker threads; the exec::par_map unfold stage)."""

    def __init__(self, meshes: Iterable[Tuple[np.ndarray, np.ndarray]], weld_tol=1e-12, angle_min_deg=180.0,
                 angle_mix=1.0, threads=0):
        pos, faces, po, fo = [], [], [0], [0]
        for p, f in meshes:
            pos.append(np.ascontiguousarray(p, np.float64).reshape(-1, 3))
            faces.append(np.ascontiguousarray(f, np.uint32).reshape(-1, 3))
            po.append(po[-1] + len(pos[-1]))
            fo.append(fo[-1] + len(faces[-1]))
        P = np.ascontiguousarray(np.concatenate(pos))
        F = np.ascontiguousarray(np.concatenate(faces))
        po = np.asarray(po, np.int64)
        fo = np.asarray(fo, np.int64)
        self.h = C.c_void_p()
        _lib.check(lib().ruv_lscm_batch_build(len(po) - 1, _p(P), _p(po), _p(F), _p(fo), weld_tol, angle_min_deg,
                                              angle_mix, threads, C.byref(self.h)))
        sz = np.zeros(4, np.int64)
        _lib.check(lib().ruv_lscm_batch_sizes(self.h, _p(sz)))
        self.nchart, self.nrows, self.nnz, self.nv_total = (int(x) for x in sz)

    def arrays(self, csr=False):
        n, c, z = self.nrows, self.nchart, self.nnz
        out = dict(rhs=np.zeros(2 * n), chart_row_off=np.zeros(c + 1, np.int64), chart_mesh=np.zeros(c, np.int32),
                   chart_nv=np.zeros(c, np.int64), chart_npins=np.zeros(c, np.int64))
        if csr:
            out.update(row_ptr=np.zeros(n + 1, np.int64), cols=np.zeros(z, np.int32), vals=np.zeros(z))
        _lib.check(lib().ruv_lscm_batch_arrays(self.h, _p(out.get("row_ptr")), _p(out.get("cols")),
                                               _p(out.get("vals")), _p(out["rhs"]), _p(out["chart_row_off"]),
                                               _p(out["chart_mesh"]), _p(out["chart_nv"]), _p(out["chart_npins"])))
        return out

    def device_batch(self, precond=Precond.JACOBI, mixed=True) -> Batch:
        h = C.c_void_p()
        _lib.check(lib().ruv_batch_create_lscm(self.h, int(precond), 1 if mixed else 0, C.byref(h)))
        return Batch(h, self.nrows, self.nchart, 2)

    def close(self):
        if self.h:
            lib().ruv_lscm_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
