"""The IC(0) factorisation on the device (factor_ic0, precond.rs:145-219): the factor the level-scheduled kernels
compute is the oracle's restatement of the reference to the bit, failures match the reference's None, and a plan's
values update refactors on the device to the same factor (and solution) as a new plan."""
import numpy as np
import pytest

from tests import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def csr(v, c, p):
    import researchuv_amd as ra

    return ra.CsrMatrix(np.asarray(v, np.float64), np.asarray(c, np.int32), np.asarray(p, np.int32))


def _height_field_system(orc, nx, ny, seed):
    """The interleaved LSCM free block of one jittered height-field chart (pins on its border)."""
    from researchuv_amd import meshgen

    pos, faces = meshgen.heightfield(nx, ny, seed=seed)
    pw, fw = orc.weld(pos, faces, 1e-12)
    s = orc.Charts(pw, fw, 180.0).system(0)
    n = 2 * s["nfree"]
    return s["vals"], s["cols"], s["row_ptr"][: n + 1], s["rhs"][:n]


def _cases(orc):
    yield "banded(400,3)", meshes.diag_dominant_banded(400, 3)
    yield "poisson 60^2 varying", meshes.poisson_2d(60, varying=True)[:3]
    v, c, p, _ = _height_field_system(orc, 48, 40, 5)
    yield "height field 48x40 (interleaved)", (v, c, p)


def test_device_factor_is_the_oracle_factor_to_the_bit(orc):
    from researchuv_amd import solver

    for name, (v, c, p) in _cases(orc):
        say(f"device ic0 factor: {name}")
        want = orc.ic0_factor(v, c, p)
        got = solver.ic0_factor(csr(v, c, p))
        np.testing.assert_array_equal(got["l_ptr"], want["l_ptr"])
        np.testing.assert_array_equal(got["l_cols"], want["l_cols"])
        np.testing.assert_array_equal(got["l_vals"], want["l_vals"])


def test_vanishing_pivot_fails_like_the_reference(orc):
    """factor_ic0 returns None on a vanishing diagonal (precond.rs:184-186): the device factorisation reports
    PRECOND for the same matrix."""
    from researchuv_amd import _lib, solver

    v, c, p, _ = meshes.laplacian_1d(64)
    v = np.array(v, np.float64)
    c = np.asarray(c, np.int32)
    p = np.asarray(p, np.int32)
    for k in range(p[10], p[11]):  # zero the diagonal of row 10
        if c[k] == 10:
            v[k] = 0.0
    say("device ic0 with a zero diagonal")
    assert orc.ic0_factor(v, c, p) is None
    with pytest.raises(_lib.RuvError) as e:
        solver.ic0_factor(csr(v, c, p))
    assert "PRECOND" in str(e.value)


@pytest.mark.parametrize("pc", ["IC0", "IC0_COLOR", "JACOBI"])
@pytest.mark.parametrize("mixed", [False, True])
def test_plan_update_values_matches_a_new_plan(orc, pc, mixed):
    """A plan given new values (same pattern) solves exactly as a plan created from them: the preconditioner is
    refreshed on the device from the kept structure."""
    from researchuv_amd import solver

    v, c, p, rhs = _height_field_system(orc, 40, 36, 9)
    n = len(p) - 1
    b = np.asarray(rhs, np.float64)
    rng = np.random.default_rng(3)
    # a second operator on the same pattern: symmetric positive scaling of the rows and columns
    s = rng.uniform(0.5, 2.0, n)
    rows = np.repeat(np.arange(n), np.diff(p))
    v2 = np.asarray(v, np.float64) * s[rows] * s[np.asarray(c)]
    prc = getattr(solver.Precond, pc)
    say(f"plan update_values {pc} mixed={mixed}")
    plan = solver.Plan(csr(v, c, p), 1, prc, mixed)
    x1, _ = plan.solve(b, 20000, 1e-10)
    plan.update_values(v2)
    x2, i2 = plan.solve(b, 20000, 1e-10)
    fresh = solver.Plan(csr(v2, c, p), 1, prc, mixed)
    x3, i3 = fresh.solve(b, 20000, 1e-10)
    np.testing.assert_array_equal(x2, x3)
    assert i2.iters == i3.iters
    # and the solution is that of the new operator
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    A2 = sp.csr_matrix((v2, c, p), shape=(n, n))
    xr = spla.spsolve(A2.tocsc(), b)
    assert np.abs(x2 - xr).max() <= 1e-6 * np.abs(xr).max()
    assert np.abs(x1 - x2).max() > 0  # the update changed the operator
    plan.close()
    fresh.close()


def test_plan_update_values_rejects_amg():
    from researchuv_amd import _lib, solver

    v, c, p, _ = meshes.poisson_2d(20, varying=False)
    plan = solver.Plan(csr(v, c, p), 1, solver.Precond.AMG, False)
    with pytest.raises(_lib.RuvError) as e:
        plan.update_values(np.asarray(v, np.float64) * 2.0)
    assert "INVALID" in str(e.value)
    plan.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_plan_solves_into_new_output_buffers(orc, mixed):
    """Solves of one plan into different device output arrays give the same solution in each: the iterations the
    plan replays from its captured graph work on the plan's own vectors, never on a caller's earlier buffer."""
    import torch

    from researchuv_amd import solver

    v, c, p, rhs = _height_field_system(orc, 32, 30, 4)
    b = torch.tensor(np.asarray(rhs, np.float64), device="cuda:0")
    plan = solver.Plan(csr(v, c, p), 1, solver.Precond.IC0, mixed)
    say(f"plan re-solves into new buffers mixed={mixed}")
    xs = []
    for _ in range(3):
        x = torch.full_like(b, float("nan"))
        plan.solve_dev(b.data_ptr(), x.data_ptr(), 20000, 1e-10)
        torch.cuda.synchronize()
        xs.append(x.cpu().numpy())
        del x
    plan.close()
    assert np.isfinite(xs[0]).all()
    np.testing.assert_array_equal(xs[0], xs[1])
    np.testing.assert_array_equal(xs[0], xs[2])
