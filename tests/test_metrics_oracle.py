"""Pins the oracle's restatement of the post-solve chart metrics (metrics.rs:31-227) against the reference's own
tests (recut.rs:155-169, metrics.rs:239-279), against closed forms of affine maps and rotations, and against
planted folds.  Nothing here needs a GPU."""
import numpy as np
import pytest


def planar_grid(nx, ny, z=None):
    """(nx+1)*(ny+1) vertices on [0,nx]x[0,ny] (z optional), two CCW triangles per cell."""
    xs, ys = np.meshgrid(np.arange(nx + 1, dtype=float), np.arange(ny + 1, dtype=float), indexing="xy")
    p = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size) if z is None else z(xs.ravel(), ys.ravel())], -1)
    tris = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            b, c, d = a + 1, a + nx + 1, a + nx + 2
            tris += [(a, b, d), (a, d, c)]
    return p, np.array(tris, np.uint32)


def border_ids(nx, ny):
    return np.array([j * (nx + 1) + i for j in range(ny + 1) for i in range(nx + 1)
                     if i in (0, nx) or j in (0, ny)], np.int64)


# ------------------------------------------------------------------ winding_flips
def test_winding_flips_reference_cases(orc):
    # recut.rs:155-169
    uv = np.array([[0.0, 0.0], [0.5, 0.0], [0.0, 0.5], [0.5, 0.5]])
    assert orc.winding_flips(uv, [[0, 1, 2], [1, 2, 3]]) == 1
    assert orc.winding_flips(uv, [[0, 2, 1], [1, 2, 3]]) == 0
    assert orc.winding_flips(uv, np.zeros((0, 3))) == 0


def test_winding_flips_counts_a_planted_fold(orc):
    p, t = planar_grid(6, 5)
    uv = p[:, :2].copy()
    assert orc.winding_flips(uv, t) == 0
    assert orc.winding_flips(uv * [-1.0, 1.0], t) == 0  # a uniform reflection is consistent
    t2 = t.copy()
    t2[7] = t2[7][[0, 2, 1]]  # one triangle reversed
    assert orc.winding_flips(uv, t2) == 1


# ------------------------------------------------------------------ chart_distortion
def test_identity_unfold_of_a_triangle_is_conformal(orc):
    # metrics.rs:239-257
    p = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    m = orc.chart_distortion(p, [[0, 1, 2]], p[:, :2])
    assert abs(m["conformal_mean"] - 1.0) < 1e-9 and m["flips"] == 0
    assert abs(m["area_ratio_mean"] - 1.0) < 1e-12 and m["folds"] == 0


@pytest.mark.parametrize("A", [[[2.0, 0.0], [0.0, 0.5]], [[2.0, 1.0], [0.5, 1.5]], [[0.0, -3.0], [1.0, 0.2]]])
def test_affine_map_of_a_planar_mesh(orc, A):
    # uv = A (x, y): every triangle's Jacobian is A up to a rotation of the tangent frame, so the conformal
    # ratio is A's singular-value ratio and the area ratio |det A| on every triangle
    A = np.array(A)
    p, t = planar_grid(5, 4)
    uv = p[:, :2] @ A.T
    s = np.linalg.svd(A, compute_uv=False)
    m = orc.chart_distortion(p, t, uv)
    assert m["conformal_mean"] == pytest.approx(s[0] / s[1], rel=1e-12)
    assert m["conformal_max"] == pytest.approx(s[0] / s[1], rel=1e-12)
    assert m["area_ratio_mean"] == pytest.approx(abs(np.linalg.det(A)), rel=1e-12)
    assert m["flips"] == 0
    # a reflecting map (det < 0) is uniformly reversed, not folded
    assert m["folds"] == 0


def test_isometric_unfold_of_a_tilted_plane(orc):
    # the plane z = 0.3 x - 0.2 y unrolled onto an orthonormal basis of itself: conformal, area-preserving
    p, t = planar_grid(4, 4, z=lambda x, y: 0.3 * x - 0.2 * y)
    n = np.array([-0.3, 0.2, 1.0]) / np.linalg.norm([-0.3, 0.2, 1.0])
    e1 = np.array([1.0, 0.0, 0.3]) / np.linalg.norm([1.0, 0.0, 0.3])
    e2 = np.cross(n, e1)
    uv = np.stack([p @ e1, p @ e2], -1)
    m = orc.chart_distortion(p, t, uv)
    assert m["conformal_mean"] == pytest.approx(1.0, abs=1e-12)
    assert m["area_ratio_mean"] == pytest.approx(1.0, abs=1e-12)


def test_degenerate_triangles_are_skipped(orc):
    p, t = planar_grid(3, 3)
    uv = p[:, :2] * [2.0, 1.0]
    m0 = orc.chart_distortion(p, t, uv)
    p2 = np.vstack([p, [[10.0, 0.0, 0.0], [11.0, 0.0, 0.0], [12.0, 0.0, 0.0]]])  # collinear: zero normal
    uv2 = np.vstack([uv, [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]])
    t2 = np.vstack([t, [[len(p), len(p) + 1, len(p) + 2]]])
    m = orc.chart_distortion(p2, t2, uv2)
    assert m["conformal_mean"] == m0["conformal_mean"] and m["area_ratio_mean"] == m0["area_ratio_mean"]
    assert m["flips"] == 0


# ------------------------------------------------------------------ rectangularize
def test_rectangularize_unit_square(orc):
    # metrics.rs:258-279
    uv = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    q, ext = orc.rectangularize(uv, [0, 1, 2, 3])
    assert abs(ext[0] - 1.0) < 1e-9 and abs(ext[1] - 1.0) < 1e-9
    for u, v in q:
        assert min(abs(u), abs(1 - u), abs(v), abs(1 - v)) < 1e-9


@pytest.mark.parametrize("k", [0, 10, 25, 40])
def test_rectangle_rotated_by_a_ladder_angle_comes_back(orc, k):
    # a 3 x 1 rectangle turned by minus the ladder's k-th angle: the ladder's k-th rotation undoes it, the
    # border-box diagonal is smallest there, and the extent is the rectangle's
    p, _ = planar_grid(6, 2)
    uv = p[:, :2] * [0.5, 0.5]
    ang = np.pi * k / 71
    R = np.array([[np.cos(ang), np.sin(ang)], [-np.sin(ang), np.cos(ang)]])
    tilted = uv @ R  # row vectors: R^T applied, the inverse of the ladder's rotation
    b = border_ids(6, 2)
    q, ext = orc.rectangularize(tilted, b)
    assert ext == pytest.approx([3.0, 1.0], abs=1e-12)
    assert q.min() >= -1e-12 and q.max() <= 1 + 1e-12
    inner = np.setdiff1d(np.arange(len(uv)), b)
    assert q[inner] == pytest.approx((uv[inner] - uv.min(0)) / [3.0, 1.0], abs=1e-12)
    for u, v in q[b]:
        assert min(u, 1 - u, v, 1 - v) < 1e-12


def test_rectangularize_without_border_only_normalises(orc):
    rng = np.random.default_rng(3)
    uv = rng.normal(size=(50, 2)) * [3.0, 0.5] + [7.0, -2.0]
    q, ext = orc.rectangularize(uv, np.zeros(0, np.int64))
    assert ext == pytest.approx(uv.max(0) - uv.min(0), rel=0, abs=0)
    assert q.min(0) == pytest.approx([0.0, 0.0], abs=0) and q.max(0) == pytest.approx([1.0, 1.0], abs=1e-15)
