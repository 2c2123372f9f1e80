"""GPU parity of the device PCG against the CPU oracle, through the C ABI.

Mirrors the reference's device tests (crates/researchuv-gpu/src/solver.rs:767-1109,
precond.rs:887-927) and its end-to-end acceptance test (lscm.rs:733-760:
max |uv_gpu - uv_cpu| / max |uv_cpu| < 1e-6 on grid(32,32))."""
import numpy as np
import pytest

from tests import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def gpu():
    import researchuv_amd as ra

    s = ra.GpuSolver.global_()
    assert s is not None, "no HIP device: the GPU tests need an MI355X"
    return s


def csr(v, c, p):
    import researchuv_amd as ra

    return ra.CsrMatrix(v, c, p)


def uv_rel_err(a, b):
    """lscm.rs:749-758: max |a - b| / max |b| (b the CPU solution)."""
    return np.linalg.norm(a - b, axis=1).max() / max(np.linalg.norm(b, axis=1).max(), 1e-30)


# ------------------------------------------------------------------ solver.rs tests
def test_cg_matches_the_direct_solution(gpu):
    say("laplacian_1d n=4096 cg_solve tol 1e-12")
    v, c, p, b = meshes.laplacian_1d(4096)
    x = gpu.cg_solve(csr(v, c, p), b, 20000, 1e-12)
    assert x is not None, gpu.last_error
    assert np.abs(x - 1.0).max() < 1e-8


def test_large_well_conditioned_system_converges(gpu):
    say("banded n=500000 tol 1e-10")
    n, band = 500_000, 3
    i = np.arange(n)
    # solver.rs:789-816: row-sorted CSR, columns i-3..i+3, -0.25/d off the diagonal, 4 on it
    offs = np.arange(-band, band + 1)
    C = i[:, None] + offs[None, :]
    V = np.where(offs == 0, 4.0, -0.25 / np.maximum(np.abs(offs), 1))
    V = np.broadcast_to(V, C.shape).copy()
    mask = (C >= 0) & (C < n)
    cols = C[mask].astype(np.int32)
    vals = V[mask]
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum(mask.sum(axis=1))
    b = (i % 97) * 0.01 + 1.0
    x = gpu.cg_solve(csr(vals, cols, ptr), b, 10_000, 1e-10)
    assert x is not None, gpu.last_error
    for r in (0, n // 2, n - 1):
        s = sum(vals[k] * x[cols[k]] for k in range(ptr[r], ptr[r + 1]))
        assert abs(s - b[r]) < 1e-6


def test_preconditioning_reduces_iterations(gpu):
    import researchuv_amd as ra

    say("poisson 400^2 none/jacobi/ic0 tol 1e-10")
    v, c, p, b = meshes.poisson_2d(400)
    a = csr(v, c, p)
    n = a.rows()
    xn, itn = gpu.cg_solve_precond(a, b, 100_000, 1e-10, ra.Precond.NONE)
    xj, itj = gpu.cg_solve_precond(a, b, 100_000, 1e-10, ra.Precond.JACOBI)
    xi, iti = gpu.cg_solve_precond(a, b, 100_000, 1e-10, ra.Precond.IC0)
    for x in (xn, xj, xi):
        for i in (0, n // 2, n - 1):
            assert abs(x[i] - xn[i]) < 1e-5
    say(f"iterations none={itn} jacobi={itj} ic0={iti}")
    assert itj <= itn and iti < itn


def test_colored_ic0_solves_the_elongated_system(gpu):
    import researchuv_amd as ra

    say("laplacian_1d n=20000 none vs ic0color")
    v, c, p, b = meshes.laplacian_1d(20_000)
    a = csr(v, c, p)
    xn, itn = gpu.cg_solve_precond(a, b, 200_000, 1e-10, ra.Precond.NONE)
    xc, itc = gpu.cg_solve_precond(a, b, 200_000, 1e-10, ra.Precond.IC0_COLOR)
    assert np.abs(xc - xn).max() < 1e-5
    assert itc < itn


def test_mismatched_dimensions_are_rejected(gpu):
    v, c, p, b = meshes.laplacian_1d(8)
    assert gpu.cg_solve(csr(v, c, p), b[:4], 10, 1e-10) is None
    assert "BAD_DIMS" in gpu.last_error


def test_failure_modes_map_to_none(gpu):
    import researchuv_amd as ra

    say("breakdown / max_iter / nan")
    # indefinite: p.Ap == 0 on the first step
    a = csr([1.0, -1.0], [0, 1], [0, 1, 2])
    assert gpu.cg_solve_precond(a, [1.0, 1.0], 10, 1e-12, ra.Precond.NONE) is None
    assert "BREAKDOWN" in gpu.last_error
    v, c, p, b = meshes.laplacian_1d(64)
    assert gpu.cg_solve_precond(csr(v, c, p), b, 1, 1e-12, ra.Precond.JACOBI) is None
    assert "NOT_CONVERGED" in gpu.last_error
    bn = b.copy()
    bn[3] = np.nan
    assert gpu.cg_solve_precond(csr(v, c, p), bn, 100, 1e-12, ra.Precond.JACOBI) is None
    # zero diagonal: Jacobi setup fails (precond.rs:596-608)
    a0 = csr([0.0, 1.0, 1.0, 2.0], [0, 1, 0, 1], [0, 2, 4])
    assert gpu.cg_solve_precond(a0, [1.0, 1.0], 10, 1e-12, ra.Precond.JACOBI) is None
    assert "PRECOND" in gpu.last_error


def test_ic0_host_solve_matches_device_apply(gpu, orc):
    say("ic0 apply banded(400,3)")
    v, c, p = meshes.diag_dominant_banded(400, 3)
    n = len(p) - 1
    r = np.array([((i * 7919) % 101) / 101.0 - 0.5 for i in range(n)])
    z_host = orc.ic0_apply(orc.ic0_factor(v, c, p), r)
    z_dev = gpu.apply_ic0_for_test(csr(v, c, p), r)
    assert z_dev is not None, gpu.last_error
    assert np.abs(z_dev - z_host).max() < 1e-9


# ------------------------------------------------------------------ LSCM parity (lscm.rs:733-760)
def lscm_case(orc, mesh, angle=30.0, chart=0):
    pw, fw = orc.weld(*mesh, 1e-12)
    ch = orc.Charts(pw, fw, angle)
    return ch, ch.system(chart), ch.unfold(chart)[0]


def uv_from(sys_, x):
    uv = sys_["p2"].copy()
    uv[sys_["free_ids"]] = x.reshape(-1, 2)
    uv[sys_["pin_ids"]] = sys_["pin_uv"]
    return uv


@pytest.mark.parametrize("pc", ["JACOBI", "IC0", "NONE", "IC0_COLOR"])
def test_lscm_grid32_gpu_matches_cpu_lu(gpu, orc, pc):
    import researchuv_amd as ra

    say(f"lscm grid(32,32) precond {pc}")
    ch, s, uv_cpu = lscm_case(orc, orc.grid(32, 32))
    a = csr(s["vals"], s["cols"], s["row_ptr"])
    n = a.rows()
    # gpu_solve (lscm.rs:500-501): max_iter = max(100, 1024) * 16, tol 1e-12
    r = gpu.cg_solve_precond(a, s["rhs"][:n], 1024 * 16, 1e-12, getattr(ra.Precond, pc))
    assert r is not None, gpu.last_error
    uv = uv_from(s, r[0])
    err = uv_rel_err(uv, uv_cpu)
    say(f"  iters={r[1]} uv_err={err:.3e}")
    assert err < 1e-6


def test_lscm_cube6_plain_ic0_all_charts(gpu, orc):
    """cube(6): 6 charts of 50 unknowns, FP64 PCG + IC(0) at 1e-12 with the 1024 gate bypassed."""
    import researchuv_amd as ra

    pw, fw = orc.weld(*orc.cube(6), 1e-12)
    ch = orc.Charts(pw, fw, 30.0)
    for i in range(6):
        s = ch.system(i)
        uv_cpu, _ = ch.unfold(i)
        a = csr(s["vals"], s["cols"], s["row_ptr"])
        r = gpu.cg_solve_precond(a, s["rhs"][: a.rows()], 1024 * 16, 1e-12, ra.Precond.IC0)
        assert r is not None, gpu.last_error
        assert uv_rel_err(uv_from(s, r[0]), uv_cpu) < 1e-9


def test_lscm_ic0_apply_on_interleaved_system_matches_oracle(gpu, orc):
    """The device streams A = L (x) I2 as L against two right-hand sides; the IC(0) operator it applies must
    equal the reference's IC(0) of the full interleaved A."""
    _, s, _ = lscm_case(orc, orc.grid(16, 16))
    n = len(s["row_ptr"]) - 1
    r = np.sin(np.arange(n) * 0.37)
    z_host = orc.ic0_apply(orc.ic0_factor(s["vals"], s["cols"], s["row_ptr"]), r)
    z_dev = gpu.apply_ic0_for_test(csr(s["vals"], s["cols"], s["row_ptr"]), r)
    assert z_dev is not None, gpu.last_error
    assert np.abs(z_dev - z_host).max() < 1e-9 * np.abs(z_host).max()


@pytest.mark.parametrize("pc", ["JACOBI", "IC0"])
def test_mixed_precision_reaches_1e8_true_residual(gpu, orc, pc):
    import researchuv_amd as ra

    say(f"mixed grid(64,64) {pc}")
    ch, s, uv_cpu = lscm_case(orc, orc.grid(64, 64))
    a = csr(s["vals"], s["cols"], s["row_ptr"])
    n = a.rows()
    b = s["rhs"][:n]
    r = gpu.cg_solve_mixed(a, b, 200_000, 1e-8, getattr(ra.Precond, pc))
    assert r is not None, gpu.last_error
    x, info = r
    import scipy.sparse as sp

    As = sp.csr_matrix((s["vals"], s["cols"], s["row_ptr"]), shape=(n, n))
    true = np.linalg.norm(b - As @ x) / np.linalg.norm(b)
    say(f"  inner iters={info.iters} refinements={info.refinements} true={true:.3e} reported={info.true_rel_resid:.3e}")
    assert true <= 1e-8 and abs(true - info.true_rel_resid) <= 1e-3 * true + 1e-15
    assert info.refinements >= 1
    assert uv_rel_err(uv_from(s, x), uv_cpu) < 1e-6


def test_plan_device_buffers_roundtrip(gpu, orc):
    import torch

    import researchuv_amd as ra

    say("plan solve_dev grid(32)")
    _, s, uv_cpu = lscm_case(orc, orc.grid(32, 32))
    a = csr(s["vals"], s["cols"], s["row_ptr"])
    n = a.rows()
    # the plan takes L directly (rhs_per_row = 2): even rows of A on even columns
    lv, lc, lp = [], [], [0]
    for i in range(0, n, 2):
        for k in range(s["row_ptr"][i], s["row_ptr"][i + 1]):
            lc.append(s["cols"][k] // 2)
            lv.append(s["vals"][k])
        lp.append(len(lv))
    plan = ra.Plan(csr(lv, lc, lp), 2, ra.Precond.JACOBI, mixed=True)
    b = torch.tensor(s["rhs"][:n], dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    torch.cuda.synchronize()  # the plan runs on its own stream
    for _ in range(2):  # repeated solves on one plan
        info = plan.solve_dev(b.data_ptr(), x.data_ptr(), 100_000, 1e-8)
        torch.cuda.synchronize()
        assert info.true_rel_resid <= 1e-8
        assert uv_rel_err(uv_from(s, x.cpu().numpy()), uv_cpu) < 1e-6
    plan.close()


# ------------------------------------------------------------------ AMG (amg.rs, solver.rs:937-1102)
def test_amg_beats_jacobi_on_the_poisson_system(gpu):
    import researchuv_amd as ra

    say("amg vs jacobi poisson 400^2")
    v, c, p, b = meshes.poisson_2d(400)
    a = csr(v, c, p)
    n = a.rows()
    xj, itj = gpu.cg_solve_precond(a, b, 100_000, 1e-10, ra.Precond.JACOBI)
    xa, ita = gpu.cg_solve_precond(a, b, 100_000, 1e-10, ra.Precond.AMG)
    xs, its = gpu.cg_solve_precond(a, b, 100_000, 1e-10, ra.Precond.AMG_SA)
    for x in (xa, xs):
        for i in (0, n // 2, n - 1):
            assert abs(x[i] - xj[i]) < 1e-4
    say(f"  jacobi={itj} amg={ita} amgsa={its}")
    assert ita < itj and its < itj


def test_smoothed_amg_on_the_elongated_system(gpu):
    import researchuv_amd as ra

    say("amgsa on 1-D laplacian n=200000")
    v, c, p, b = meshes.laplacian_1d(200_000)
    a = csr(v, c, p)
    xs, its = gpu.cg_solve_precond(a, b, 200_000, 1e-10, ra.Precond.AMG_SA)
    assert np.abs(xs - 1.0).max() < 1e-5
    say(f"  amgsa iterations {its}")
    assert its < 500


@pytest.mark.parametrize("mixed", [False, True])
def test_lscm_grid64_amgsa_matches_cpu_lu(gpu, orc, mixed):
    import researchuv_amd as ra

    say(f"lscm grid(64,64) amgsa mixed={mixed}")
    ch, s, uv_cpu = lscm_case(orc, orc.grid(64, 64))
    a = csr(s["vals"], s["cols"], s["row_ptr"])
    n = a.rows()
    if mixed:
        r = gpu.cg_solve_mixed(a, s["rhs"][:n], 100_000, 1e-9, ra.Precond.AMG_SA)
    else:
        r = gpu.cg_solve_precond(a, s["rhs"][:n], 1024 * 16, 1e-12, ra.Precond.AMG_SA)
    assert r is not None, gpu.last_error
    err = uv_rel_err(uv_from(s, r[0]), uv_cpu)
    say(f"  uv_err={err:.3e}")
    assert err < 1e-6


@pytest.mark.parametrize("pc", ["JACOBI", "AMG_SA", "IC0"])
def test_graph_replay_matches_direct_launches(gpu, orc, pc, monkeypatch):
    """The PCG iterations replay a captured HIP graph; every replay (several solves on one plan, with eager
    work and synchronisations in between) must give bitwise the result of direct launches."""
    import researchuv_amd as ra

    say(f"graph vs direct {pc}")
    _, s, _ = lscm_case(orc, orc.grid(48, 48))
    n = len(s["row_ptr"]) - 1
    lv, lc, lp = [], [], [0]
    for i in range(0, n, 2):
        for k in range(s["row_ptr"][i], s["row_ptr"][i + 1]):
            lc.append(s["cols"][k] // 2)
            lv.append(s["vals"][k])
        lp.append(len(lv))
    Lm = csr(lv, lc, lp)
    b = s["rhs"][:n]
    monkeypatch.setenv("RUV_NO_GRAPH", "1")
    direct = ra.Plan(Lm, 2, getattr(ra.Precond, pc), mixed=True)
    monkeypatch.delenv("RUV_NO_GRAPH")
    graph = ra.Plan(Lm, 2, getattr(ra.Precond, pc), mixed=True)
    x0, i0 = direct.solve(b, 100_000, 1e-10)
    for rep in range(3):
        x1, i1 = graph.solve(b * (1.0 + rep), 100_000, 1e-10)
        xd, idd = direct.solve(b * (1.0 + rep), 100_000, 1e-10)
        assert np.array_equal(x1, xd), rep
        assert i1.iters == idd.iters
    direct.close()
    graph.close()


@pytest.mark.parametrize("pc", ["IC0", "IC0_COLOR"])
def test_triangular_sweep_wide_levels_match_single_workgroup(gpu, orc, pc, monkeypatch):
    """A triangular-sweep level of many rows runs as its own many-workgroup launch, the narrow levels in one
    single-workgroup launch per run; every row keeps its arithmetic, so the solves are bitwise those of the
    all-single-workgroup sweeps (levels forced wide from 32 rows here)."""
    import researchuv_amd as ra

    say(f"triangular sweeps, wide levels vs single workgroup, {pc}")
    _, s, _ = lscm_case(orc, orc.grid(64, 64))
    n = len(s["row_ptr"]) - 1
    lv, lc, lp = [], [], [0]
    for i in range(0, n, 2):
        for k in range(s["row_ptr"][i], s["row_ptr"][i + 1]):
            lc.append(s["cols"][k] // 2)
            lv.append(s["vals"][k])
        lp.append(len(lv))
    Lm = csr(lv, lc, lp)
    b = s["rhs"][:n]
    out = {}
    for wide in ("32", "1000000000"):
        monkeypatch.setenv("RUV_TRI_WIDE", wide)
        for mixed in (False, True):
            plan = ra.Plan(Lm, 2, getattr(ra.Precond, pc), mixed=mixed)
            x, info = plan.solve(b, 100_000, 1e-10)
            plan.close()
            out[(wide, mixed)] = (x, info.iters)
    for mixed in (False, True):
        xw, iw = out[("32", mixed)]
        xs, is_ = out[("1000000000", mixed)]
        assert np.array_equal(xw, xs) and iw == is_, mixed
