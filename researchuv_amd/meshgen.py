"""Seeded synthetic meshes for the LSCM workloads (numpy; deterministic for a given seed on any machine).

Triangulation follows meshgen::grid (crates/researchuv-unwrap/src/meshgen.rs:135-155): quad (a, b, c, d) with
a = j*stride + i, b = a + 1, c = a + stride, d = c + 1 gives triangles [a, b, d] and [a, d, c].

heightfield(nx, ny): vertices on [0, X] x [0, Y] with square cells of side h, x/y jittered by U(-jitter*h,
+jitter*h), and z = sum_{k<8} amp * sin(2*pi*f_k*(x cos t_k + y sin t_k) + phi_k) evaluated at the jittered
x, y, with f_k ~ U[1, 16], t_k, phi_k ~ U[0, 2*pi) -- the single-large-chart workload is nx = ny = 512 on the
unit square.  stretched_grid: planar grid with geometric column spacing and a fraction of flipped quad
diagonals (the ill-conditioned workload).  patch_batch: uneven open grid patches with log-uniform vertex
counts and aspect ratios in [1, 8].
"""
from __future__ import annotations

import numpy as np


def grid_faces(nx: int, ny: int, flip: np.ndarray | None = None) -> np.ndarray:
    """Faces of an nx x ny vertex grid (meshgen.rs:144-153); `flip` (ny-1, nx-1) bool swaps the diagonal of a
    quad to [a, b, c], [b, d, c]."""
    j, i = np.meshgrid(np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    a = (j * nx + i).astype(np.uint32)
    b = a + 1
    c = a + nx
    d = c + 1
    t0 = np.stack([a, b, d], -1)
    t1 = np.stack([a, d, c], -1)
    if flip is not None:
        f = flip[..., None]
        t0 = np.where(f, np.stack([a, b, c], -1), t0)
        t1 = np.where(f, np.stack([b, d, c], -1), t1)
    return np.stack([t0, t1], axis=2).reshape(-1, 3).astype(np.uint32)


def heightfield(nx: int, ny: int, seed: int = 0, amp: float = 0.1, jitter: float = 0.2, size: float | None = None,
                n_waves: int = 8, fmin: float = 1.0, fmax: float = 16.0):
    """(positions (nx*ny, 3) f64, faces u32).  size: extent of the longer side (default 1.0)."""
    rng = np.random.default_rng(seed)
    size = 1.0 if size is None else size
    h = size / (max(nx, ny) - 1)
    f = rng.uniform(fmin, fmax, n_waves)
    th = rng.uniform(0.0, 2.0 * np.pi, n_waves)
    ph = rng.uniform(0.0, 2.0 * np.pi, n_waves)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    x = ii * h + rng.uniform(-jitter * h, jitter * h, (ny, nx))
    y = jj * h + rng.uniform(-jitter * h, jitter * h, (ny, nx))
    z = np.zeros_like(x)
    for k in range(n_waves):
        z += amp * np.sin(2.0 * np.pi * f[k] * (x * np.cos(th[k]) + y * np.sin(th[k])) + ph[k])
    pos = np.stack([x, y, z], -1).reshape(-1, 3)
    return np.ascontiguousarray(pos), grid_faces(nx, ny)


def stretched_grid(n: int, stretch: float = 100.0, flip_frac: float = 0.05, seed: int = 0):
    """Planar n x n grid: column spacing r^i with r = stretch^(1/(n-2)) (1 -> stretch), row spacing 1, and a
    seeded fraction of quads with the other diagonal."""
    rng = np.random.default_rng(seed)
    r = stretch ** (1.0 / max(n - 2, 1))
    dx = r ** np.arange(n - 1)
    xs = np.concatenate([[0.0], np.cumsum(dx)])
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pos = np.stack([xs[ii], jj.astype(np.float64), np.zeros((n, n))], -1).reshape(-1, 3)
    flip = rng.random((n - 1, n - 1)) < flip_frac
    return np.ascontiguousarray(pos), grid_faces(n, n, flip)


def patch_sizes(count: int, vmin: int = 64, vmax: int = 65536, amin: float = 1.0, amax: float = 8.0, seed: int = 0):
    """(nx, ny) per patch: vertex count log-uniform in [vmin, vmax], aspect nx/ny uniform in [amin, amax]."""
    rng = np.random.default_rng(seed)
    nv = np.rint(np.exp(rng.uniform(np.log(vmin), np.log(vmax), count)))
    asp = rng.uniform(amin, amax, count)
    ny = np.maximum(2, np.rint(np.sqrt(nv / asp))).astype(int)
    nx = np.maximum(2, np.rint(nv / ny)).astype(int)
    return np.stack([nx, ny], -1)


def patch_batch(count: int, seed: int = 0, **kw):
    """Yields (positions, faces) of `count` uneven heightfield patches (each its own mesh and chart)."""
    sizes = patch_sizes(count, seed=seed, **kw)
    for k, (nx, ny) in enumerate(sizes):
        yield heightfield(int(nx), int(ny), seed=seed * 100003 + k + 1)
