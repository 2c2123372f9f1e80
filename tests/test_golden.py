"""The committed golden vectors (tests/golden/, written by tests/golden/make_golden.py).

Expected solutions come from SuperLU and dense numpy solves, independent of the oracle's LU; the assembled
systems and pins are regression vectors with checksums.  CPU tests check the oracle and the product's host
assembly against them; the GPU tests solve the same systems through the C ABI."""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CHARTS = {
    "cube6_chart0": lambda orc: orc.cube(6),
    "grid32_chart0": lambda orc: orc.grid(32, 32),
}


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def interleave(l_vals, l_cols, l_ptr):
    """A = L (x) I2 in (u0, v0, u1, v1, ...) order, as Sparse::to_csr lays it out (lscm.rs:616-629)."""
    L = sp.csr_matrix((l_vals, l_cols, l_ptr), shape=(len(l_ptr) - 1, len(l_ptr) - 1))
    A = sp.kron(L, sp.identity(2), format="csr")
    A.sort_indices()
    return A


@pytest.mark.parametrize("name", sorted(CHARTS))
def test_oracle_reproduces_golden_system_and_pins(orc, name):
    g = load(name)
    pins = json.load(open(os.path.join(GOLD, "pins.json")))[name]
    pos, faces = CHARTS[name](orc)
    pw, fw = orc.weld(pos, faces, 1e-12)
    s = orc.Charts(pw, fw, 30.0).system(0)
    assert sha(s["pin_ids"], s["pin_uv"]) == pins["sha256"]
    assert sha(s["vals"], s["cols"], s["row_ptr"], s["rhs"]) == pins["system_sha256"]
    for k in ("vals", "cols", "row_ptr", "rhs", "pin_ids", "pin_uv", "free_ids", "p2"):
        np.testing.assert_array_equal(s[k], g[k])


@pytest.mark.parametrize("name", sorted(CHARTS))
def test_oracle_lu_matches_superlu_golden_solution(orc, name):
    g = load(name)
    n = len(g["row_ptr"]) - 1
    x = orc.csr_sparse_solve(g["vals"], g["cols"], g["row_ptr"], g["rhs"][:n])
    assert x is not None
    assert np.abs(x - g["x"]).max() <= 1e-12 * max(1.0, np.abs(g["x"]).max())


@pytest.mark.parametrize("name", sorted(CHARTS))
def test_product_host_assembly_matches_golden(orc, name):
    from researchuv_amd import lscm

    g = load(name)
    pos, faces = CHARTS[name](orc)
    pw, fw = lscm.weld(pos, faces, 1e-12)
    s = lscm.segment(pw, fw, 30.0).system(0)
    np.testing.assert_array_equal(s["pin_ids"], g["pin_ids"])
    np.testing.assert_array_equal(s["pin_uv"], g["pin_uv"])
    np.testing.assert_array_equal(s["free_ids"], g["free_ids"])
    A = interleave(s["l_vals"], s["l_cols"], s["l_ptr"])
    np.testing.assert_array_equal(A.indptr, g["row_ptr"])
    np.testing.assert_array_equal(A.indices, g["cols"])
    np.testing.assert_array_equal(A.data, g["vals"])
    np.testing.assert_array_equal(s["rhs"], g["rhs"][: 2 * s["nfree"]])


def test_cube_golden_solution_is_the_isometric_face(orc):
    """Closed form: a cube face unfolds to its own affine image, so the free uv are the face's grid
    positions (1/6 spacing, centred) up to the rectangle's symmetry."""
    g = load("cube6_chart0")
    uv = g["x"].reshape(-1, 2)
    k = np.round(uv * 6.0)
    assert np.abs(uv * 6.0 - k).max() < 1e-6  # the pins sit on the perimeter to ~1e-8 (border_targets)
    assert sorted(map(tuple, k.astype(int))) == sorted((i, j) for i in range(-2, 3) for j in range(-2, 3))


def test_oracle_solvers_match_small_dense_answers(orc):
    z = load("spd_small")
    for k in range(int(z["count"])):
        v, c, p, b, x = (z[f"m{k}_{s}"] for s in ("vals", "cols", "row_ptr", "b", "x"))
        xs = orc.csr_sparse_solve(v, c, p, b)
        assert np.abs(xs - x).max() <= 1e-10 * max(1.0, np.abs(x).max()), k
        xc = orc.cg_solve(v, c, p, b, 10 * len(b))
        assert np.abs(xc - x).max() <= 1e-8 * max(1.0, np.abs(x).max()), k


# ------------------------------------------------------------------ GPU, through the C ABI
@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CHARTS))
@pytest.mark.parametrize("pc", ["JACOBI", "IC0", "IC0_COLOR", "AMG", "AMG_SA"])
def test_gpu_solves_golden_chart(name, pc):
    from researchuv_amd import solver

    print(f"[gpu-test] golden {name} {pc}", flush=True)
    g = load(name)
    n = len(g["row_ptr"]) - 1
    A = solver.CsrMatrix(g["vals"], g["cols"], g["row_ptr"])
    gs = solver.GpuSolver.new()
    assert gs is not None
    r = gs.cg_solve_precond(A, g["rhs"][:n], 20000, 1e-12, getattr(solver.Precond, pc))
    assert r is not None, gs.last_error
    x, _ = r
    assert np.abs(x - g["x"]).max() <= 1e-6 * max(1.0, np.abs(g["x"]).max())
    r = gs.cg_solve_mixed(A, g["rhs"][:n], 20000, 1e-10, getattr(solver.Precond, pc))
    assert r is not None, gs.last_error
    assert np.abs(r[0] - g["x"]).max() <= 1e-6 * max(1.0, np.abs(g["x"]).max())


@pytest.mark.gpu
def test_gpu_solves_small_golden_systems():
    from researchuv_amd import solver

    print("[gpu-test] golden small SPD systems", flush=True)
    z = load("spd_small")
    gs = solver.GpuSolver.new()
    for k in range(int(z["count"])):
        v, c, p, b, x = (z[f"m{k}_{s}"] for s in ("vals", "cols", "row_ptr", "b", "x"))
        for pc in (solver.Precond.NONE, solver.Precond.JACOBI, solver.Precond.IC0):
            r = gs.cg_solve_precond(solver.CsrMatrix(v, c, p), b, 10 * len(b), 1e-13, pc)
            assert r is not None, gs.last_error
            assert np.abs(r[0] - x).max() <= 1e-9 * max(1.0, np.abs(x).max()), (k, pc)
