"""ctypes binding of the CPU oracle (oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load this module.  The researchuv_amd package never imports
it.  Each wrapper mirrors one reference function; see oracle.c for the
file:line each restatement follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.orc_cube.restype = C.c_int64
        L.orc_cube.argtypes = [C.c_int, P, P]
        L.orc_grid.argtypes = [C.c_int, C.c_int, P, P]
        L.orc_weld.restype = C.c_int
        L.orc_weld.argtypes = [P, C.c_int, P, C.c_int, C.c_double, P, P, P]
        L.orc_segment.restype = P
        L.orc_segment.argtypes = [P, C.c_int, P, C.c_int, C.c_double]
        L.orc_charts_count.argtypes = [P]
        L.orc_charts_ncut.argtypes = [P]
        L.orc_charts_cut.argtypes = [P, P]
        L.orc_chart_sizes.argtypes = [P, C.c_int, P]
        L.orc_chart_get.argtypes = [P, C.c_int, P, P, P, P, P]
        L.orc_chart_loop_len.argtypes = [P, C.c_int, C.c_int]
        L.orc_chart_loop.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_charts_free.argtypes = [P]
        L.orc_eig3.argtypes = [P, P, P]
        L.orc_tangent_frame.argtypes = [P, C.c_int, P]
        L.orc_rect_point.argtypes = [C.c_double, C.c_double, C.c_double, P]
        L.orc_anchor_ids.restype = C.c_int
        L.orc_anchor_ids.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_sparse_solve.restype = C.c_int
        L.orc_sparse_solve.argtypes = [C.c_int32, C.c_int64, P, P, P, P, P]
        L.orc_chart_system.restype = C.c_int
        L.orc_chart_system.argtypes = [P, C.c_int, P, C.c_double, C.c_double, P, P, P, P, P, P, P, P, P]
        L.orc_unfold_chart.restype = C.c_int
        L.orc_unfold_chart.argtypes = [P, C.c_int, P, C.c_int, C.c_double, C.c_double, P, P]
        L.orc_par_unfold.restype = C.c_int
        L.orc_par_unfold.argtypes = [P, P, C.c_int, C.c_double, C.c_double, P, P]
        L.orc_par_unfold_iters.restype = C.c_int
        L.orc_par_unfold_iters.argtypes = [P, P, C.c_int, C.c_int, C.c_double, C.c_double, P, P]
        L.orc_cg_solve.restype = C.c_int
        L.orc_cg_solve.argtypes = [C.c_int32, P, P, P, P, C.c_int64, P]
        L.orc_ic0_factor.restype = C.c_int64
        L.orc_ic0_factor.argtypes = [C.c_int32, P, P, P, P, P, P]
        L.orc_tri_levels.restype = C.c_int32
        L.orc_tri_levels.argtypes = [C.c_int32, P, P, C.c_int, P]
        L.orc_ic0_apply.argtypes = [C.c_int32, P, P, P, P, P, P, P, P]
        L.orc_transpose_lower.argtypes = [C.c_int32, P, P, P, P, P, P]
        L.orc_coloring.argtypes = [C.c_int32, P, P, P]
        L.orc_winding_flips.restype = C.c_int64
        L.orc_winding_flips.argtypes = [P, P, C.c_int64]
        L.orc_chart_distortion.argtypes = [P, P, C.c_int64, P, P, P]
        L.orc_rectangularize.argtypes = [P, C.c_int64, P, C.c_int64, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ------------------------------------------------------------------ meshgen / weld
def cube(n):
    """meshgen.rs:12-54 -> (positions (N,3) f64, faces (F,3) u32)."""
    k = lib().orc_cube(n, None, None)
    np_, nf = k >> 32, k & 0xFFFFFFFF
    pos = np.zeros((np_, 3), np.float64)
    faces = np.zeros((nf, 3), np.uint32)
    lib().orc_cube(n, _p(pos), _p(faces))
    return pos, faces


def grid(w, h):
    """meshgen.rs:135-155."""
    pos = np.zeros(((w + 1) * (h + 1), 3), np.float64)
    faces = np.zeros((2 * w * h, 3), np.uint32)
    lib().orc_grid(w, h, _p(pos), _p(faces))
    return pos, faces


def weld(pos, faces, tol):
    """weld.rs:19-57 -> (positions, faces)."""
    pos = np.ascontiguousarray(pos, np.float64)
    faces = np.ascontiguousarray(faces, np.uint32)
    po = np.zeros_like(pos)
    fo = np.zeros_like(faces)
    nf = C.c_int(0)
    nw = lib().orc_weld(_p(pos), len(pos), _p(faces), len(faces), tol, _p(po), _p(fo), C.byref(nf))
    return po[:nw].copy(), fo[: nf.value].copy()


# ------------------------------------------------------------------ segment
class Charts:
    """Owns an orc_charts handle (segment.rs:130-223 output)."""

    def __init__(self, pos, faces, angle_min_deg):
        self.pos = np.ascontiguousarray(pos, np.float64)
        self.faces = np.ascontiguousarray(faces, np.uint32)
        self.h = lib().orc_segment(_p(self.pos), len(self.pos), _p(self.faces), len(self.faces), angle_min_deg)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_charts_free(self.h)
            self.h = None

    def __len__(self):
        return lib().orc_charts_count(self.h)

    def cut(self):
        n = lib().orc_charts_ncut(self.h)
        out = np.zeros((n, 2), np.uint32)
        lib().orc_charts_cut(self.h, _p(out))
        return out

    def chart(self, i):
        sz = np.zeros(4, np.int32)
        lib().orc_chart_sizes(self.h, i, _p(sz))
        nf, nv, nb, nl = (int(x) for x in sz)
        face_ids = np.zeros(nf, np.uint32)
        vids = np.zeros(nv, np.uint32)
        tris = np.zeros((nf, 3), np.uint32)
        border = np.zeros(nb, np.uint32)
        area = C.c_double(0)
        lib().orc_chart_get(self.h, i, _p(face_ids), _p(vids), _p(tris), _p(border), C.byref(area))
        loops = []
        for l in range(nl):
            L = lib().orc_chart_loop_len(self.h, i, l)
            arr = np.zeros(L, np.uint32)
            lib().orc_chart_loop(self.h, i, l, _p(arr))
            loops.append(arr)
        return dict(face_ids=face_ids, vertex_ids=vids, tris=tris, border=border, border_loops=loops,
                    area3d=area.value)

    def system(self, i, angle_mix=1.0, mix_w=1.0):
        """Pins, free map and the assembled free block (lscm.rs:512-676; Sparse::to_csr)."""
        sizes = np.zeros(4, np.int64)
        L = lib()
        L.orc_chart_system(self.h, i, _p(self.pos), angle_mix, mix_w, _p(sizes), None, None, None, None, None,
                           None, None, None)
        nv, nfree, npins, nnz = (int(x) for x in sizes)
        p2 = np.zeros((nv, 2))
        pin_ids = np.zeros(npins, np.int32)
        pin_uv = np.zeros((npins, 2))
        free_ids = np.zeros(nfree, np.int32)
        vals = np.zeros(nnz)
        cols = np.zeros(nnz, np.int32)
        row_ptr = np.zeros(2 * nfree + 1, np.int32)
        rhs = np.zeros(2 * nfree + 1)
        L.orc_chart_system(self.h, i, _p(self.pos), angle_mix, mix_w, _p(sizes), _p(p2), _p(pin_ids), _p(pin_uv),
                           _p(free_ids), _p(vals), _p(cols), _p(row_ptr), _p(rhs))
        return dict(nv=nv, nfree=nfree, p2=p2, pin_ids=pin_ids, pin_uv=pin_uv, free_ids=free_ids, vals=vals,
                    cols=cols, row_ptr=row_ptr, rhs=rhs)

    def unfold(self, i, max_iter=0, angle_mix=1.0, mix_w=1.0):
        """unfold_chart with the CPU sparse-LU backend (lscm.rs:512-727) -> (uv (nv,2), stats)."""
        sz = np.zeros(4, np.int32)
        lib().orc_chart_sizes(self.h, i, _p(sz))
        uv = np.zeros((int(sz[1]), 2))
        st = np.zeros(4)
        lib().orc_unfold_chart(self.h, i, _p(self.pos), max_iter, angle_mix, mix_w, _p(uv), _p(st))
        return uv, dict(iters=int(st[0]), err_accum=st[1], target=st[2], area3d=st[3])

    def par_unfold(self, workers, angle_mix=1.0, mix_w=1.0, max_iter=0):
        """Worker-thread path over every chart (pipeline.rs:171-194, exec.rs:28-60).  max_iter 0 runs the
        driver's own iteration count; 1 one factor+solve per chart."""
        n = len(self)
        sizes = []
        for i in range(n):
            sz = np.zeros(4, np.int32)
            lib().orc_chart_sizes(self.h, i, _p(sz))
            sizes.append(int(sz[1]))
        off = np.zeros(n + 1, np.int64)
        off[1:] = np.cumsum(sizes)
        uv = np.zeros((int(off[-1]), 2))
        lib().orc_par_unfold_iters(self.h, _p(self.pos), workers, max_iter, angle_mix, mix_w, _p(off), _p(uv))
        return [uv[off[i]:off[i + 1]] for i in range(n)]


# ------------------------------------------------------------------ small pieces
def eig3(m):
    m = np.ascontiguousarray(m, np.float64).reshape(9)
    w = np.zeros(3)
    v = np.zeros(9)
    lib().orc_eig3(_p(m), _p(w), _p(v))
    return w, v.reshape(3, 3)


def tangent_frame(p3):
    p3 = np.ascontiguousarray(p3, np.float64)
    out = np.zeros((len(p3), 2))
    lib().orc_tangent_frame(_p(p3), len(p3), _p(out))
    return out


def rect_point(f, a, b):
    out = np.zeros(2)
    lib().orc_rect_point(f, a, b, _p(out))
    return out


def anchor_ids(p3, n):
    p3 = np.ascontiguousarray(p3, np.float64)
    out = np.zeros(max(n, 1), np.int64)
    k = lib().orc_anchor_ids(_p(p3), len(p3), n, _p(out))
    return out[:k].copy()


def sparse_solve(n, rows, cols, vals, b):
    """Sparse::add in COO order then Sparse::solve (sparse.rs:47-52, 219-265). None on failure."""
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    rc = lib().orc_sparse_solve(n, len(vals), _p(rows), _p(cols), _p(vals), _p(b), _p(x))
    return None if rc != 0 else x


def csr_sparse_solve(vals, cols, row_ptr, b):
    n = len(row_ptr) - 1
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(row_ptr))
    return sparse_solve(n, rows, cols, vals, b)


def cg_solve(vals, cols, row_ptr, rhs, max_iter):
    """lscm::cg_solve (lscm.rs:417-459)."""
    n = len(row_ptr) - 1
    x = np.zeros(n)
    rc = lib().orc_cg_solve(n, _p(np.ascontiguousarray(vals, np.float64)), _p(np.ascontiguousarray(cols, np.int32)),
                            _p(np.ascontiguousarray(row_ptr, np.int32)), _p(np.ascontiguousarray(rhs, np.float64)),
                            max_iter, _p(x))
    return None if rc != 0 else x


def ic0_factor(vals, cols, row_ptr):
    """factor_ic0 + transpose_lower + tri_levels (precond.rs:145-290). None on a failed pivot."""
    vals = np.ascontiguousarray(vals, np.float64)
    cols = np.ascontiguousarray(cols, np.int32)
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    n = len(row_ptr) - 1
    L = lib()
    nnz = L.orc_ic0_factor(n, _p(vals), _p(cols), _p(row_ptr), None, None, None)
    lv = np.zeros(nnz)
    lc = np.zeros(nnz, np.int32)
    lp = np.zeros(n + 1, np.int32)
    if L.orc_ic0_factor(n, _p(vals), _p(cols), _p(row_ptr), _p(lv), _p(lc), _p(lp)) < 0:
        return None
    tv = np.zeros(nnz)
    tc = np.zeros(nnz, np.int32)
    tp = np.zeros(n + 1, np.int32)
    L.orc_transpose_lower(n, _p(lv), _p(lc), _p(lp), _p(tv), _p(tc), _p(tp))
    fl = np.zeros(n, np.int32)
    bl = np.zeros(n, np.int32)
    nf = L.orc_tri_levels(n, _p(lc), _p(lp), 0, _p(fl))
    nb = L.orc_tri_levels(n, _p(tc), _p(tp), 1, _p(bl))
    return dict(l_vals=lv, l_cols=lc, l_ptr=lp, lt_vals=tv, lt_cols=tc, lt_ptr=tp, fwd_level=fl, bwd_level=bl,
                n_fwd=int(nf), n_bwd=int(nb))


def ic0_apply(f, r):
    r = np.ascontiguousarray(r, np.float64)
    z = np.zeros_like(r)
    lib().orc_ic0_apply(len(r), _p(f["l_vals"]), _p(f["l_cols"]), _p(f["l_ptr"]), _p(f["lt_vals"]),
                        _p(f["lt_cols"]), _p(f["lt_ptr"]), _p(r), _p(z))
    return z


def coloring(cols, row_ptr):
    n = len(row_ptr) - 1
    out = np.zeros(n, np.uint32)
    lib().orc_coloring(n, _p(np.ascontiguousarray(cols, np.int32)), _p(np.ascontiguousarray(row_ptr, np.int32)),
                       _p(out))
    return out


def lscm_loss_frozen(p3, tris, pin_ids, pin_uv, target):
    """E + |uv - target|^2 with frozen pins, solved by the sparse LU -> (loss, uv) (None when the LU fails)."""
    L = lib()
    if not getattr(L, "_loss_ready", False):
        P = C.c_void_p
        L.orc_lscm_loss_frozen.restype = C.c_double
        L.orc_lscm_loss_frozen.argtypes = [P, C.c_int, P, C.c_int, P, P, C.c_int, P, P, P]
        L._loss_ready = True
    p3 = np.ascontiguousarray(p3, np.float64)
    tris = np.ascontiguousarray(tris, np.uint32)
    pin_ids = np.ascontiguousarray(pin_ids, np.int32)
    pin_uv = np.ascontiguousarray(pin_uv, np.float64)
    target = np.ascontiguousarray(target, np.float64)
    uv = np.zeros((len(p3), 2))
    ok = C.c_int(0)
    loss = L.orc_lscm_loss_frozen(_p(p3), len(p3), _p(tris), len(tris), _p(pin_ids), _p(pin_uv), len(pin_ids),
                                  _p(target), _p(uv), C.byref(ok))
    return (loss, uv) if ok.value else None


def _vjp_setup(L):
    if not getattr(L, "_vjp_ready", False):
        P = C.c_void_p
        L.orc_lscm_vjp_frozen.restype = C.c_int
        L.orc_lscm_vjp_frozen.argtypes = [P, C.c_int, P, C.c_int, P, P, C.c_int, P, C.c_double, P, P]
        L.orc_lscm_grad_scatter.restype = None
        L.orc_lscm_grad_scatter.argtypes = [P, C.c_int, P, C.c_int, P, P, C.c_double, P]
        L._vjp_ready = True
    return L


def lscm_vjp_frozen(p3, tris, pin_ids, pin_uv, g_uv, energy_weight=1.0):
    """dJ/dp for J = energy_weight * E(uv(p), p) + <g_uv, uv(p)> (frozen pins) by two sparse-LU solves with the
    same matrix (L^T = L) -> (uv, grad (nv, 3)); None when an LU solve fails."""
    L = _vjp_setup(lib())
    p3 = np.ascontiguousarray(p3, np.float64)
    tris = np.ascontiguousarray(tris, np.uint32)
    pin_ids = np.ascontiguousarray(pin_ids, np.int32)
    pin_uv = np.ascontiguousarray(pin_uv, np.float64)
    g_uv = np.ascontiguousarray(g_uv, np.float64)
    uv = np.zeros((len(p3), 2))
    grad = np.zeros((len(p3), 3))
    rc = L.orc_lscm_vjp_frozen(_p(p3), len(p3), _p(tris), len(tris), _p(pin_ids), _p(pin_uv), len(pin_ids), _p(g_uv),
                               float(energy_weight), _p(uv), _p(grad))
    return (uv, grad) if rc == 0 else None


def lscm_grad_scatter(p3, tris, uv, lam, energy_weight=1.0):
    """The assembly half of the VJP given the solution uv and the adjoint lam (zero at pins) -> grad (nv, 3)."""
    L = _vjp_setup(lib())
    p3 = np.ascontiguousarray(p3, np.float64)
    tris = np.ascontiguousarray(tris, np.uint32)
    uv = np.ascontiguousarray(uv, np.float64)
    lam = np.ascontiguousarray(lam, np.float64)
    grad = np.zeros((len(p3), 3))
    L.orc_lscm_grad_scatter(_p(p3), len(p3), _p(tris), len(tris), _p(uv), _p(lam), float(energy_weight), _p(grad))
    return grad


# ------------------------------------------------------------------ post-solve chart metrics
def winding_flips(uv, tris):
    """metrics.rs:39-56"""
    uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    return int(lib().orc_winding_flips(_p(uv), _p(tris), len(tris)))


def chart_distortion(p3, tris, uv):
    """metrics.rs:73-141 on chart-local positions / triangles / uv -> dict of the ChartMetrics fields."""
    p3 = np.ascontiguousarray(p3, np.float64).reshape(-1, 3)
    tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    out = np.zeros(3)
    cnt = np.zeros(2, np.int64)
    lib().orc_chart_distortion(_p(p3), _p(tris), len(tris), _p(uv), _p(out), _p(cnt))
    return dict(conformal_mean=float(out[0]), conformal_max=float(out[1]), area_ratio_mean=float(out[2]),
                flips=int(cnt[0]), folds=int(cnt[1]))


def rectangularize(uv, bnd):
    """metrics.rs:147-227 -> (q (nv, 2), ext (2,))"""
    uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    bnd = np.ascontiguousarray(bnd, np.int64).reshape(-1)
    q = np.zeros_like(uv)
    ext = np.zeros(2)
    lib().orc_rectangularize(_p(uv), len(uv), _p(bnd), len(bnd), _p(q), _p(ext))
    return q, ext
