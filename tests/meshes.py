"""Test-side mesh generators in numpy, restating meshgen.rs for the fixtures the
C oracle does not carry (sphere, torus, annulus, cylinder) plus small helper
matrices used by the reference's own solver tests."""
import numpy as np


def uv_sphere(nu, nv):
    """meshgen.rs:59-86 (both ends included; weld to close it)."""
    du = 2.0 * np.pi / (nu - 1)
    dv = np.pi / (nv - 1)
    p = []
    for j in range(nv):
        v = dv * j
        for i in range(nu):
            u = du * i
            p.append((np.sin(v) * np.cos(u), np.sin(v) * np.sin(u), np.cos(v)))
    f = []
    for i in range(nu):
        for j in range(nv - 1):
            a = j * nu + i
            b = j * nu + (i + 1) % nu
            c = a + nu
            d = b + nu
            f.append((a, c, d))
            f.append((a, d, b))
    return np.array(p, np.float64), np.array(f, np.uint32)


def torus(R, r, nu, nvt):
    """meshgen.rs:91-114."""
    p = []
    for j in range(nvt):
        v = 2.0 * np.pi * j / nvt
        for i in range(nu):
            u = 2.0 * np.pi * i / nu
            cr = R + r * np.cos(v)
            p.append((cr * np.cos(u), cr * np.sin(u), r * np.sin(v)))
    f = []
    for i in range(nu):
        for j in range(nvt):
            a = j * nu + i
            b = j * nu + (i + 1) % nu
            jn = (j + 1) % nvt
            c = jn * nu + i
            d = jn * nu + (i + 1) % nu
            f.append((a, b, d))
            f.append((a, d, c))
    return np.array(p, np.float64), np.array(f, np.uint32)


def torus_annulus(R, r, nu, nvt):
    """meshgen.rs:120-130."""
    p, f = torus(R, r, nu, nvt)
    jc = nvt // 2
    keep = [(k // 2) % nvt != jc for k in range(len(f))]
    return p, f[np.array(keep)]


def cylinder(r, height, nu, nv):
    """meshgen.rs:159-181."""
    p = []
    for j in range(nv + 1):
        z = height * j / max(nv, 1)
        for i in range(nu + 1):
            t = 2.0 * np.pi * i / max(nu, 1)
            p.append((r * np.cos(t), r * np.sin(t), z))
    f = []
    stride = nu + 1
    for j in range(nv):
        for i in range(nu):
            a = j * stride + i
            b = a + 1
            c = a + stride
            d = c + 1
            f.append((a, b, d))
            f.append((a, d, c))
    return np.array(p, np.float64), np.array(f, np.uint32)


def pinched_grid():
    """A flat 3 x 3 quad grid (vertex i + 4 j at (i, j, 0)) without the centre quad and the corner quad
    (2, 2): the hole and the notch meet at vertex 10 = (2, 2), which has four boundary edges.  The seven
    remaining quads form one edge-connected ring."""
    p = np.array([(i, j, 0.0) for j in range(4) for i in range(4)], np.float64)
    f = []
    for j in range(3):
        for i in range(3):
            if (i, j) in ((1, 1), (2, 2)):
                continue
            a = j * 4 + i
            f += [(a, a + 1, a + 5), (a, a + 5, a + 4)]
    return p, np.array(f, np.uint32)


def pinched_grid_tie():
    """The pinched grid with its top row moved one unit to -x: vertex 14 sits at (1, 3), so from the
    pinch vertex 10 = (2, 2) the neighbours 9 = (1, 2) and 14 lie at the same distance along -x."""
    p, f = pinched_grid()
    p[12:, 0] -= 1.0
    return p, f


def laplacian_1d(n):
    """solver.rs:743-765: tridiagonal (-1, 2, -1); x = 1 solves b = e_0 + e_{n-1}."""
    vals, cols, ptr = [], [], [0]
    for i in range(n):
        if i > 0:
            vals.append(-1.0)
            cols.append(i - 1)
        vals.append(2.0)
        cols.append(i)
        if i + 1 < n:
            vals.append(-1.0)
            cols.append(i + 1)
        ptr.append(len(vals))
    b = np.zeros(n)
    b[0] = 1.0
    b[-1] = 1.0
    return np.array(vals), np.array(cols, np.int32), np.array(ptr, np.int32), b


def diag_dominant_banded(n, band):
    """precond.rs:683-706."""
    vals, cols, ptr = [], [], [0]
    for i in range(n):
        for d in range(1, band + 1):
            if i >= d:
                vals.append(-0.25 / d)
                cols.append(i - d)
        vals.append(4.0 + (i % 11) * 0.75)
        cols.append(i)
        for d in range(1, band + 1):
            if i + d < n:
                vals.append(-0.25 / d)
                cols.append(i + d)
        ptr.append(len(vals))
    # rows as built are not column-sorted for the lower part (i-1, i-2, ...); sort each row
    vals = np.array(vals)
    cols = np.array(cols, np.int32)
    ptr = np.array(ptr, np.int32)
    for i in range(n):
        s, e = ptr[i], ptr[i + 1]
        o = np.argsort(cols[s:e], kind="stable")
        cols[s:e] = cols[s:e][o]
        vals[s:e] = vals[s:e][o]
    return vals, cols, ptr


def poisson_2d(side, varying=True):
    """solver.rs:837-862: 5-point (4 / -1), diagonal 4.05 on every 11th row."""
    n = side * side
    vals, cols, ptr = [], [], [0]
    for r in range(side):
        for c in range(side):
            i = r * side + c
            ent = []
            for rr, cc in ((r - 1, c), (r, c - 1), (r + 1, c), (r, c + 1)):
                if 0 <= rr < side and 0 <= cc < side:
                    ent.append((rr * side + cc, -1.0))
            ent.append((i, 4.05 if (varying and i % 11 == 0) else 4.0))
            ent.sort()
            for cc, v in ent:
                cols.append(cc)
                vals.append(v)
            ptr.append(len(vals))
    b = np.array([(i % 97) * 0.01 + 1.0 for i in range(n)])
    return np.array(vals), np.array(cols, np.int32), np.array(ptr, np.int32), b


def to_dense(vals, cols, ptr):
    n = len(ptr) - 1
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(ptr[i], ptr[i + 1]):
            A[i, cols[k]] += vals[k]
    return A
