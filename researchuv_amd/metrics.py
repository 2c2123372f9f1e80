"""Post-solve chart metrics on the device: `chart_distortion`, `winding_flips`, `rectangularize`.

Mirrors crates/researchuv-unwrap/src/metrics.rs (`ChartMetrics` :15-29, `winding_flips` :39-56,
`chart_distortion` :73-141, `rectangularize` :147-227): the same names, arguments and results, one chart or a
batch of charts per call (one workgroup per chart, ruv_chart_metrics / ruv_rectangularize).  A chart is given
by its vertices in chart-local order -- `p3`, the mesh positions of its vertex_ids, and `uv`, the unfold --
and its triangles as chart-local corner triples in mesh face order; `bnd` lists its border vertices (distinct
chart-local ids).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

from . import _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _setup(L):
    if getattr(L, "_ruv_metrics_ready", False):
        return L
    P = C.c_void_p
    L.ruv_chart_metrics.restype = C.c_int
    L.ruv_chart_metrics.argtypes = [C.c_int32, P, P, P, P, P, P, P]
    L.ruv_rectangularize.restype = C.c_int
    L.ruv_rectangularize.argtypes = [C.c_int32, P, P, P, P, P, P]
    L.ruv_chart_metrics_dev.restype = C.c_int
    L.ruv_chart_metrics_dev.argtypes = [C.c_int32, P, P, P, P, P, P, P]
    L.ruv_chart_metrics_dev_noalloc.restype = C.c_int
    L.ruv_chart_metrics_dev_noalloc.argtypes = [C.c_int32, P, P, P, P, P, P, P]
    L.ruv_rectangularize_dev.restype = C.c_int
    L.ruv_rectangularize_dev.argtypes = [C.c_int32, P, P, P, P, P, P]
    L._ruv_metrics_ready = True
    return L


def lib():
    return _setup(_lib.lib())


@dataclass
class ChartMetrics:
    """metrics.rs:15-29"""
    conformal_mean: float
    conformal_max: float
    area_ratio_mean: float
    flips: int
    folds: int


def _offsets(sizes) -> np.ndarray:
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(np.asarray(sizes, np.int64))
    return off


def chart_distortion_batch(charts: Sequence[Tuple[np.ndarray, np.ndarray, np.ndarray]]) -> List[ChartMetrics]:
    """charts: (p3 (nv,3), tris (nt,3) chart-local, uv (nv,2)) per chart -> ChartMetrics per chart."""
    n = len(charts)
    p3 = [np.ascontiguousarray(c[0], np.float64).reshape(-1, 3) for c in charts]
    tris = [np.ascontiguousarray(c[1], np.uint32).reshape(-1, 3) for c in charts]
    uv = [np.ascontiguousarray(c[2], np.float64).reshape(-1, 2) for c in charts]
    for a, b in zip(p3, uv):
        if len(a) != len(b):
            raise ValueError("p3 and uv of a chart must have one row per chart vertex")
    v_off = _offsets([len(a) for a in p3])
    t_off = _offsets([len(t) for t in tris])
    P3 = np.concatenate(p3) if n else np.zeros((0, 3))
    UV = np.concatenate(uv) if n else np.zeros((0, 2))
    T = np.concatenate(tris) if n else np.zeros((0, 3), np.uint32)
    met = np.zeros((n, 3))
    cnt = np.zeros((n, 2), np.int64)
    _lib.check(lib().ruv_chart_metrics(n, _p(v_off), _p(P3), _p(UV), _p(t_off), _p(T), _p(met), _p(cnt)))
    return [ChartMetrics(float(m[0]), float(m[1]), float(m[2]), int(k[0]), int(k[1])) for m, k in zip(met, cnt)]


def chart_distortion(p3, tris, uv) -> ChartMetrics:
    """metrics.rs:73-141 for one chart."""
    return chart_distortion_batch([(p3, tris, uv)])[0]


def winding_flips(uv, tris) -> int:
    """metrics.rs:39-56: triangles whose UV winding disagrees with the chart's signed-area majority."""
    uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    p3 = np.zeros((len(uv), 3))  # no 3-D term is read by the fold count
    return chart_distortion(p3, tris, uv).folds


def rectangularize_batch(charts: Sequence[Tuple[np.ndarray, np.ndarray]]) -> List[Tuple[np.ndarray, np.ndarray]]:
    """charts: (uv (nv,2), bnd (nb,) distinct chart-local border ids) -> (q (nv,2), ext (2,)) per chart."""
    n = len(charts)
    uv = [np.ascontiguousarray(c[0], np.float64).reshape(-1, 2) for c in charts]
    bnd = [np.ascontiguousarray(c[1], np.int64).reshape(-1) for c in charts]
    v_off = _offsets([len(a) for a in uv])
    b_off = _offsets([len(b) for b in bnd])
    UV = np.concatenate(uv) if n else np.zeros((0, 2))
    BN = np.concatenate(bnd) if n else np.zeros(0, np.int64)
    q = np.zeros_like(UV)
    ext = np.zeros((n, 2))
    _lib.check(lib().ruv_rectangularize(n, _p(v_off), _p(UV), _p(b_off), _p(BN), _p(q), _p(ext)))
    return [(q[v_off[i]:v_off[i + 1]], ext[i]) for i in range(n)]


def rectangularize(uv, bnd) -> Tuple[np.ndarray, np.ndarray]:
    """metrics.rs:147-227 for one chart -> (q, ext)."""
    return rectangularize_batch([(uv, bnd)])[0]
