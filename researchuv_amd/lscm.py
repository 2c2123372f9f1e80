"""Python mirror of the unfold path's chart-level interface (crates/researchuv-unwrap: weld.rs, segment.rs,
lscm.rs), backed by the C++ host library and the device solver.

Names follow the reference: `weld`, `segment`, `Chart`, `UnfoldOptions`, `SolverBackend`, `unfold_chart`.
Only the GPU backend is implemented here; the CPU backend is the reference's own sparse LU, which a caller
keeps as its fallback (lscm.rs:677-681) -- `unfold_chart` raises `RuvError` where the Rust returned None from
`gpu_solve`.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from .solver import Precond


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _setup(L):
    if getattr(L, "_ruv_lscm_ready", False):
        return L
    P = C.c_void_p
    L.ruv_weld.restype = C.c_int
    L.ruv_weld.argtypes = [P, C.c_int64, P, C.c_int64, C.c_double, P, P, P, P]
    L.ruv_segment.restype = C.c_int
    L.ruv_segment.argtypes = [P, C.c_int64, P, C.c_int64, C.c_double, P]
    L.ruv_charts_count.restype = C.c_int64
    L.ruv_charts_count.argtypes = [P]
    L.ruv_chart_sizes.restype = C.c_int
    L.ruv_chart_sizes.argtypes = [P, C.c_int64, P, P]
    L.ruv_chart_arrays.restype = C.c_int
    L.ruv_chart_arrays.argtypes = [P, C.c_int64, P, P, P, P, P, P]
    L.ruv_charts_destroy.restype = None
    L.ruv_charts_destroy.argtypes = [P]
    L.ruv_lscm_system.restype = C.c_int
    L.ruv_lscm_system.argtypes = [P, P, C.c_int64, C.c_double, P, P, P, P, P, P, P, P, P]
    L.ruv_lscm_system_device.restype = C.c_int
    L.ruv_lscm_system_device.argtypes = [P, P, C.c_int64, C.c_double, P, P, P, P, P]
    L.ruv_unfold_chart.restype = C.c_int
    L.ruv_unfold_chart.argtypes = [P, P, C.c_int64, P, P, P, P]
    L._ruv_lscm_ready = True
    return L


def lib():
    return _setup(_lib.lib())


class SolverBackend(enum.Enum):
    """lscm.rs:57-66"""

    CPU = 0
    GPU = 1


@dataclass
class UnfoldOptions:
    """lscm.rs:75-101 plus the GPU-path settings (preconditioner from RESEARCHUV_GPU_PRECOND, mode, gate)."""

    max_iter: int = 100
    angle_mix: float = 1.0
    keep_metric: bool = False
    target_area: float = 0.0
    mix_w: float = 1.0
    solver: SolverBackend = SolverBackend.GPU
    precond: Precond = field(default_factory=lambda: Precond.from_env_name(os.environ.get("RESEARCHUV_GPU_PRECOND")))
    mixed: bool = False
    min_vars: int = 1024  # GPU_MIN_VARS (lscm.rs:462)
    tol: float = 1e-12    # lscm.rs:501
    device_assembly: bool = True

    def c_opts(self):
        return _lib.UnfoldOpts(self.max_iter, self.angle_mix, self.mix_w, int(self.precond), 1 if self.mixed else 0,
                               self.min_vars, self.tol, 1 if self.device_assembly else 0)


def weld(positions, faces, tol=1e-12):
    """weld.rs:19-57 -> (positions (V,3), faces (F,3))."""
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
    po = np.zeros_like(pos)
    fo = np.zeros_like(f)
    nv = C.c_int64(0)
    nf = C.c_int64(0)
    _lib.check(lib().ruv_weld(_p(pos), len(pos), _p(f), len(f), tol, _p(po), _p(fo), C.byref(nv), C.byref(nf)))
    return po[: nv.value].copy(), fo[: nf.value].copy()


@dataclass
class Chart:
    """segment::Chart (segment.rs:32-46)"""

    face_ids: np.ndarray
    vertex_ids: np.ndarray
    tris: np.ndarray
    border: np.ndarray
    border_loops: List[np.ndarray]
    area3d: float

    def local_of(self, v) -> Optional[int]:
        i = int(np.searchsorted(self.vertex_ids, v))
        return i if i < len(self.vertex_ids) and self.vertex_ids[i] == v else None

    def is_borderless(self) -> bool:
        return len(self.border_loops) == 0


class Segmentation:
    """The charts of a welded mesh (segment::segment, segment.rs:130-223); owns the native handle."""

    def __init__(self, positions, faces, angle_min_deg=30.0):
        self.positions = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        self.h = C.c_void_p()
        _lib.check(lib().ruv_segment(_p(self.positions), len(self.positions), _p(self.faces), len(self.faces),
                                     angle_min_deg, C.byref(self.h)))

    def __len__(self):
        return int(lib().ruv_charts_count(self.h))

    def __del__(self):
        try:
            if self.h:
                lib().ruv_charts_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def n_cut(self) -> int:
        if len(self) == 0:
            return 0
        sz = np.zeros(6, np.int64)
        _lib.check(lib().ruv_chart_sizes(self.h, 0, _p(sz), None))
        return int(sz[5])

    def chart(self, i) -> Chart:
        sz = np.zeros(6, np.int64)
        area = C.c_double(0)
        _lib.check(lib().ruv_chart_sizes(self.h, i, _p(sz), C.byref(area)))
        nf, nv, nb, nl, lt = (int(x) for x in sz[:5])
        fi = np.zeros(nf, np.uint32)
        vi = np.zeros(nv, np.uint32)
        tr = np.zeros((nf, 3), np.uint32)
        bo = np.zeros(nb, np.uint32)
        lo = np.zeros(max(lt, 1), np.uint32)
        off = np.zeros(nl + 1, np.int64)
        _lib.check(lib().ruv_chart_arrays(self.h, i, _p(fi), _p(vi), _p(tr), _p(bo), _p(lo), _p(off)))
        loops = [lo[off[k]:off[k + 1]].copy() for k in range(nl)]
        return Chart(fi, vi, tr, bo, loops, area.value)

    def system(self, i, angle_mix=1.0):
        """Pins, free map, L and the interleaved rhs of chart i (lscm.rs:512-666)."""
        L = lib()
        sz = np.zeros(5, np.int64)
        _lib.check(L.ruv_lscm_system(_p(self.positions), self.h, i, angle_mix, _p(sz), None, None, None, None, None,
                                     None, None, None))
        nv, nfree, npins, nnz, _ = (int(x) for x in sz)
        out = dict(nv=nv, nfree=nfree, edges=int(sz[4]), p2=np.zeros((nv, 2)), pin_ids=np.zeros(npins, np.int32),
                   pin_uv=np.zeros((npins, 2)), free_ids=np.zeros(nfree, np.int32), l_vals=np.zeros(nnz),
                   l_cols=np.zeros(nnz, np.int32), l_ptr=np.zeros(nfree + 1, np.int32), rhs=np.zeros(2 * nfree))
        _lib.check(L.ruv_lscm_system(_p(self.positions), self.h, i, angle_mix, _p(sz), _p(out["p2"]),
                                     _p(out["pin_ids"]), _p(out["pin_uv"]), _p(out["free_ids"]), _p(out["l_vals"]),
                                     _p(out["l_cols"]), _p(out["l_ptr"]), _p(out["rhs"])))
        return out

    def system_device(self, i, angle_mix=1.0):
        """L and rhs of chart i assembled on the device (bitwise the host assembly)."""
        L = lib()
        sz = np.zeros(5, np.int64)
        _lib.check(L.ruv_lscm_system_device(_p(self.positions), self.h, i, angle_mix, _p(sz), None, None, None, None))
        nv, nfree, npins, nnz, _ = (int(x) for x in sz)
        out = dict(nv=nv, nfree=nfree, edges=int(sz[4]), l_vals=np.zeros(nnz), l_cols=np.zeros(nnz, np.int32),
                   l_ptr=np.zeros(nfree + 1, np.int32), rhs=np.zeros(2 * nfree))
        _lib.check(L.ruv_lscm_system_device(_p(self.positions), self.h, i, angle_mix, _p(sz), _p(out["l_vals"]),
                                            _p(out["l_cols"]), _p(out["l_ptr"]), _p(out["rhs"])))
        return out

    def unfold_chart(self, i, opts: Optional[UnfoldOptions] = None):
        """lscm::unfold_chart (lscm.rs:512-727), GPU backend.  Returns (uv (nv,2), stats, SolveInfo)."""
        opts = opts or UnfoldOptions()
        if opts.keep_metric:
            raise _lib.RuvError(-8, "keep_metric systems are non-symmetric: CPU only (lscm.rs:667-681)")
        sz = np.zeros(6, np.int64)
        _lib.check(lib().ruv_chart_sizes(self.h, i, _p(sz), None))
        uv = np.zeros((int(sz[1]), 2))
        st = np.zeros(4)
        info = _lib.SolveInfo()
        o = opts.c_opts()
        _lib.check(lib().ruv_unfold_chart(_p(self.positions), self.h, i, C.byref(o), _p(uv), _p(st), C.byref(info)))
        return uv, dict(iters=int(st[0]), err_accum=st[1], target=st[2], area3d=st[3]), info


def segment(positions, faces, angle_min_deg=30.0) -> Segmentation:
    return Segmentation(positions, faces, angle_min_deg)


@dataclass
class ChartResult:
    """pipeline::ChartResult (pipeline.rs:180-189) of one chart of the batched unfold stage."""

    mesh: int
    vertex_ids: np.ndarray      # the welded mesh's vertex of each chart vertex (ascending)
    uv: np.ndarray              # (nv, 2), pins at their targets
    rectified: np.ndarray       # (nv, 2)
    ext: np.ndarray             # (2,)
    conformal_mean: float
    conformal_max: float
    area_ratio_mean: float
    flips: int
    folds: int                  # recut::chart_folds' measure (recut.rs:71-80)
    iters: int                  # UnfoldResult (lscm.rs:104-116)
    err_accum: float
    target: float
    area3d: float


def unfold_batch(meshes, opts: Optional[UnfoldOptions] = None, weld_tol=1e-12, angle_min_deg=30.0, threads=0):
    """The unfold stage of pipeline::run (pipeline.rs:171-194) over many meshes in one call: every chart's solve
    in one batched device solve, the driver iterations per chart on host threads, rectangularize and
    chart_distortion on the device.  meshes: iterable of (positions (V,3), faces (F,3)).  -> ([ChartResult], info)"""
    opts = opts or UnfoldOptions(precond=Precond.AMG, mixed=True, tol=1e-8)
    L = lib()
    if not getattr(L, "_ruv_unfold_batch_ready", False):
        P = C.c_void_p
        L.ruv_unfold_batch_run.restype = C.c_int
        L.ruv_unfold_batch_run.argtypes = [C.c_int32, P, P, P, P, C.c_double, C.c_double, P, C.c_int32, P]
        L.ruv_unfold_batch_sizes.restype = C.c_int
        L.ruv_unfold_batch_sizes.argtypes = [P, P]
        L.ruv_unfold_batch_results.restype = C.c_int
        L.ruv_unfold_batch_results.argtypes = [P] * 11
        L.ruv_unfold_batch_destroy.restype = None
        L.ruv_unfold_batch_destroy.argtypes = [P]
        L._ruv_unfold_batch_ready = True
    pos, faces, po, fo = [], [], [0], [0]
    for p, f in meshes:
        pos.append(np.ascontiguousarray(p, np.float64).reshape(-1, 3))
        faces.append(np.ascontiguousarray(f, np.uint32).reshape(-1, 3))
        po.append(po[-1] + len(pos[-1]))
        fo.append(fo[-1] + len(faces[-1]))
    Pp = np.ascontiguousarray(np.concatenate(pos)) if pos else np.zeros((0, 3))
    F = np.ascontiguousarray(np.concatenate(faces)) if faces else np.zeros((0, 3), np.uint32)
    po = np.asarray(po, np.int64)
    fo = np.asarray(fo, np.int64)
    h = C.c_void_p()
    o = opts.c_opts()
    _lib.check(L.ruv_unfold_batch_run(len(po) - 1, _p(Pp), _p(po), _p(F), _p(fo), weld_tol, angle_min_deg,
                                      C.byref(o), threads, C.byref(h)))
    try:
        sz = np.zeros(3, np.int64)
        _lib.check(L.ruv_unfold_batch_sizes(h, _p(sz)))
        nc, nv = int(sz[0]), int(sz[1])
        cm = np.zeros(nc, np.int32)
        off = np.zeros(nc + 1, np.int64)
        vid = np.zeros(nv, np.uint32)
        uv, rq = np.zeros((nv, 2)), np.zeros((nv, 2))
        ext, met, cnt, st = np.zeros((nc, 2)), np.zeros((nc, 3)), np.zeros((nc, 2), np.int64), np.zeros((nc, 4))
        info = _lib.SolveInfo()
        _lib.check(L.ruv_unfold_batch_results(h, _p(cm), _p(off), _p(vid), _p(uv), _p(rq), _p(ext), _p(met), _p(cnt),
                                              _p(st), C.byref(info)))
    finally:
        L.ruv_unfold_batch_destroy(h)
    out = []
    for c in range(nc):
        s = slice(int(off[c]), int(off[c + 1]))
        out.append(ChartResult(int(cm[c]), vid[s].copy(), uv[s].copy(), rq[s].copy(), ext[c].copy(), float(met[c, 0]),
                               float(met[c, 1]), float(met[c, 2]), int(cnt[c, 0]), int(cnt[c, 1]), int(st[c, 0]),
                               float(st[c, 1]), float(st[c, 2]), float(st[c, 3])))
    return out, info
