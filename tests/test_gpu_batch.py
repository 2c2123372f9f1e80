"""GPU parity of the batched solver: every chart of a batch against the CPU oracle's sparse LU of the same
chart, plus the batch-level invariants (per-chart convergence, independence from the batch composition)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _meshes(n, seed=3):
    from researchuv_amd import meshgen

    sizes = meshgen.patch_sizes(n, vmin=64, vmax=1500, seed=seed)
    return [meshgen.heightfield(int(a), int(b), seed=seed * 1000 + k + 1) for k, (a, b) in enumerate(sizes)]


def _oracle_solutions(orc, meshes):
    out = []
    for pos, faces in meshes:
        pw, fw = orc.weld(pos, faces, 1e-12)
        ch = orc.Charts(pw, fw, 180.0)
        for i in range(len(ch)):
            s = ch.system(i)
            n = 2 * s["nfree"]
            x = orc.csr_sparse_solve(s["vals"], s["cols"], s["row_ptr"], s["rhs"][:n])
            out.append(x)
    return out


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_matches_per_chart_lu(orc, mixed):
    from researchuv_amd import batch as B

    say(f"batch of 24 uneven charts mixed={mixed}")
    meshes = _meshes(24)
    lb = B.LscmBatchBuilder(meshes, angle_min_deg=180.0, threads=4)
    a = lb.arrays()
    bt = lb.device_batch(mixed=mixed)
    x, info, it, res = bt.solve(a["rhs"], max_iter=20000, tol=1e-10 if not mixed else 1e-9)
    ref = _oracle_solutions(orc, meshes)
    assert len(ref) == lb.nchart
    off = a["chart_row_off"]
    worst = 0.0
    for c in range(lb.nchart):
        xc = x[2 * off[c]:2 * off[c + 1]].reshape(-1, 2)
        rc = ref[c].reshape(-1, 2)
        err = np.linalg.norm(xc - rc, axis=1).max() / max(np.linalg.norm(rc, axis=1).max(), 1e-30)
        worst = max(worst, err)
    say(f"  worst chart uv err {worst:.2e}, max iters {info.iters}, worst true resid {info.true_rel_resid:.2e}")
    assert worst < 1e-6
    if mixed:
        assert (res <= 1e-9).all()
    assert (it > 0).sum() > 0


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_largest_patches_at_bench_tolerance(mixed):
    """The four largest patches of the 4096-patch workload (64 227-65 404 vertices, the same seeds and shapes),
    with eight small patches beside them, solved with the device AMG at the workload's tolerance (true relative
    residual 1e-8) and checked per chart against SuperLU: the UV metric of lscm.rs:749-758 stays below 1e-6.
    A residual test bounds the UV error least on the largest charts."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from researchuv_amd import batch as B
    from researchuv_amd import meshgen
    from researchuv_amd.solver import Precond

    sizes = meshgen.patch_sizes(4096, seed=0)
    top = np.argsort(-(sizes[:, 0] * sizes[:, 1]), kind="stable")[:4]
    meshes = [meshgen.heightfield(int(sizes[k][0]), int(sizes[k][1]), seed=int(k) + 1) for k in top]
    meshes += _meshes(8, seed=9)
    say(f"largest patches {[tuple(sizes[k]) for k in top]} + 8 small, mixed={mixed}, tol 1e-8")
    lb = B.LscmBatchBuilder(meshes, angle_min_deg=180.0, threads=8)
    a = lb.arrays(csr=True)
    bt = lb.device_batch(precond=Precond.AMG, mixed=mixed)
    x, info, it, res = bt.solve(a["rhs"], max_iter=100000, tol=1e-8)
    bt.close()
    off, ptr = a["chart_row_off"], a["row_ptr"]
    worst = 0.0
    for c in range(lb.nchart):
        r0, r1 = int(off[c]), int(off[c + 1])
        L = sp.csr_matrix((a["vals"][ptr[r0]:ptr[r1]], a["cols"][ptr[r0]:ptr[r1]] - r0, ptr[r0:r1 + 1] - ptr[r0]),
                          shape=(r1 - r0, r1 - r0))
        xr = spla.splu(L.tocsc()).solve(a["rhs"][2 * r0:2 * r1].reshape(-1, 2).copy())
        xg = x[2 * r0:2 * r1].reshape(-1, 2)
        err = np.linalg.norm(xg - xr, axis=1).max() / np.linalg.norm(xr, axis=1).max()
        worst = max(worst, err)
        if c < 4:
            say(f"  chart {c}: {r1 - r0} free vertices, {int(it[c])} iters, true resid {res[c]:.2e}, uv err {err:.2e}")
    say(f"  worst uv err {worst:.2e}, worst true resid {info.true_rel_resid:.2e}")
    assert max(np.diff(off)[:4]) > 60000
    assert worst < 1e-6
    if mixed:
        assert (res <= 1e-8).all()


def test_batch_is_independent_of_composition(orc, monkeypatch):
    """Each chart runs its own PCG: solving a chart alone or inside a batch gives the same iterates (with the
    tile schedule held fixed, since the tiles decide the order of the partial sums)."""
    from researchuv_amd import batch as B

    monkeypatch.setenv("RUV_BATCH_SLICES_PER_TILE", "2")

    meshes = _meshes(6, seed=5)
    lb = B.LscmBatchBuilder(meshes, threads=2)
    a = lb.arrays()
    x_all, _, it_all, _ = lb.device_batch(mixed=False).solve(a["rhs"], tol=1e-10)
    off = a["chart_row_off"]
    for c in (0, 3, 5):
        lbc = B.LscmBatchBuilder([meshes[a["chart_mesh"][c]]], threads=1)
        ac = lbc.arrays()
        xc, _, itc, _ = lbc.device_batch(mixed=False).solve(ac["rhs"], tol=1e-10)
        assert itc[0] == it_all[c]
        np.testing.assert_array_equal(xc, x_all[2 * off[c]:2 * off[c + 1]])


def test_batch_schedules_agree(orc, monkeypatch):
    """Every tile schedule the autotuner can pick solves the same batch to the same tolerance."""
    from researchuv_amd import batch as B

    meshes = _meshes(10, seed=8)
    lb = B.LscmBatchBuilder(meshes, threads=2)
    a = lb.arrays()
    xs = []
    for spt in (1, 2, 4, 8):
        monkeypatch.setenv("RUV_BATCH_SLICES_PER_TILE", str(spt))
        say(f"schedule {spt} slices per tile")
        bt = lb.device_batch(mixed=True)
        assert bt.schedule()["rows_per_tile"] == 256 * spt
        assert bt.schedule()["col_bytes"] == 2
        x, info, it, res = bt.solve(a["rhs"], max_iter=20000, tol=1e-10)
        assert (res <= 1e-10).all()
        xs.append(x)
        bt.close()
    for x in xs[1:]:
        assert np.abs(x - xs[0]).max() <= 1e-7 * np.abs(xs[0]).max()


def test_batch_wide_columns_and_empty_chart():
    """A chart whose rows are randomly permuted has column offsets beyond 16 bits (32-bit column path); an empty
    chart between two charts keeps its own (trivial) scalars.  Checked against SuperLU per chart."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from researchuv_amd import batch as B
    from researchuv_amd import lscm, meshgen

    say("wide-column batch (permuted 190x190 chart) + empty chart")
    pos, faces = meshgen.heightfield(190, 190, seed=4)
    pw, fw = lscm.weld(pos, faces, 1e-12)
    s = lscm.segment(pw, fw, 180.0).system(0)
    n = s["nfree"]
    L = sp.csr_matrix((s["l_vals"], s["l_cols"], s["l_ptr"]), shape=(n, n))
    perm = np.random.default_rng(1).permutation(n)
    Lp = L[perm][:, perm].tocsr()
    Lp.sort_indices()
    rhs = s["rhs"].reshape(-1, 2)[perm]
    small = meshgen.heightfield(20, 20, seed=2)
    s2 = lscm.segment(*lscm.weld(*small, 1e-12), 180.0).system(0)
    L2 = sp.csr_matrix((s2["l_vals"], s2["l_cols"], s2["l_ptr"]), shape=(s2["nfree"],) * 2)
    A = sp.block_diag([Lp, L2], format="csr")
    A.sort_indices()
    off = [0, n, n, n + s2["nfree"]]  # chart 1 is empty
    b = np.concatenate([rhs.reshape(-1), s2["rhs"]])
    bt = B.Batch.from_csr(A.indptr.astype(np.int64), A.indices, A.data, off, mixed=True)
    assert bt.schedule()["col_bytes"] == 4
    x, info, it, res = bt.solve(b, max_iter=40000, tol=1e-10)
    assert it[1] == 0
    lu = spla.splu(Lp.tocsc())
    xr = np.stack([lu.solve(rhs[:, 0]), lu.solve(rhs[:, 1])], axis=1)
    xa = x[:2 * n].reshape(-1, 2)
    err = np.linalg.norm(xa - xr, axis=1).max() / np.linalg.norm(xr, axis=1).max()
    lu2 = spla.splu(L2.tocsc())
    r2 = s2["rhs"].reshape(-1, 2)
    xr2 = np.stack([lu2.solve(r2[:, 0]), lu2.solve(r2[:, 1])], axis=1)
    err2 = np.linalg.norm(x[2 * n:].reshape(-1, 2) - xr2, axis=1).max() / np.linalg.norm(xr2, axis=1).max()
    say(f"  uv err permuted chart {err:.2e}, small chart {err2:.2e}")
    assert err < 1e-6 and err2 < 1e-6


def _grid_operator(nx, ny, seed, diag_frac):
    """SPD grid operator on nx x ny rows in row-major order: couplings to the right and lower neighbours, and to
    each lower diagonal neighbour with probability diag_frac (offsets 1, nx, nx + 1, nx - 1), diagonally dominant."""
    import scipy.sparse as sp

    rng = np.random.default_rng(seed)
    r, c, v = [], [], []
    for y in range(ny):
        for x in range(nx):
            i = y * nx + x
            for dx, dy in ((1, 0), (0, 1), (1, 1), (-1, 1)):
                X, Y = x + dx, y + dy
                if not (0 <= X < nx and Y < ny):
                    continue
                if dx * dy != 0 and rng.random() >= diag_frac:
                    continue
                w = 0.5 + rng.random()
                r += [i, Y * nx + X]
                c += [Y * nx + X, i]
                v += [-w, -w]
    A = sp.csr_matrix((v, (r, c)), shape=(nx * ny,) * 2)
    A = (A + sp.diags(-np.asarray(A.sum(axis=1)).ravel() + 0.01 + 0.1 * rng.random(nx * ny))).tocsr()
    A.sort_indices()
    return A


@pytest.mark.parametrize("mixed", [True, False])
@pytest.mark.parametrize("precond", ["JACOBI", "AMG"])
def test_batch_band_form(mixed, precond, monkeypatch):
    """Charts in band form (grid operators with offsets 1, W - 1, W, W + 1: all four, some, a 3-row-high grid
    wider than a workgroup) beside charts that are not (a permuted grid, a 1-D chain, a one-row chart), every
    tile schedule: each chart against SuperLU, and the batch against the same batch with the band form off."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    mats = [_grid_operator(50, 37, 1, 1.0), _grid_operator(31, 40, 2, 0.5), _grid_operator(300, 3, 3, 1.0),
            _grid_operator(64, 64, 4, 0.0)]
    P = _grid_operator(45, 45, 5, 1.0)
    perm = np.random.default_rng(6).permutation(P.shape[0])
    mats.append(P[perm][:, perm].tocsr())
    chain = sp.diags([-np.ones(199), 2.01 * np.ones(200), -np.ones(199)], [-1, 0, 1], format="csr")
    mats += [chain, sp.csr_matrix(np.array([[3.0]]))]
    A = sp.block_diag(mats, format="csr")
    A.sort_indices()
    off = np.cumsum([0] + [m.shape[0] for m in mats])
    b = np.random.default_rng(7).standard_normal(2 * A.shape[0])
    refs = []
    for k, m in enumerate(mats):
        lu = spla.splu(sp.csc_matrix(m))
        rb = b[2 * off[k]:2 * off[k + 1]].reshape(-1, 2)
        refs.append(np.stack([lu.solve(rb[:, 0]), lu.solve(rb[:, 1])], axis=1))
    pc = getattr(Precond, precond)
    sols = {}
    for band in ("1", "0"):
        monkeypatch.setenv("RUV_BATCH_BAND", band)
        for spt in ("1", "2", "4", "8"):
            monkeypatch.setenv("RUV_BATCH_SLICES_PER_TILE", spt)
            say(f"band form {band}, {spt} slices per tile, {precond} mixed={mixed}")
            bt = B.Batch.from_csr(A.indptr.astype(np.int64), A.indices, A.data, off, precond=pc, mixed=mixed)
            assert bt.band_info()["band_charts"] == (4 if band == "1" else 0)
            x, info, it, res = bt.solve(b, max_iter=40000, tol=1e-10)
            bt.close()
            for k in range(len(mats)):
                xr = refs[k]
                err = np.linalg.norm(x[2 * off[k]:2 * off[k + 1]].reshape(-1, 2) - xr, axis=1).max()
                assert err <= 1e-6 * np.linalg.norm(xr, axis=1).max(), (k, err)
            sols[band, spt] = x
    for key, x in sols.items():
        assert np.abs(x - sols["0", "2"]).max() <= 1e-6 * np.abs(sols["0", "2"]).max(), key


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_amg_setup_from_band_form_is_bitwise_the_csr_setup(mixed, monkeypatch):
    """The finest level's matching reads band charts' rows from the band form (the same entries in the same column
    order as the setup CSR): every level of the hierarchy -- aggregate maps, charts, operator structure and values --
    equals, bit for bit, the hierarchy built from the CSR with the band form off."""
    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    say(f"AMG setup from the band form against the CSR setup, mixed={mixed}")
    meshes = _meshes(10, seed=31)
    lb = B.LscmBatchBuilder(meshes, threads=2)
    monkeypatch.setenv("RUV_BATCH_SLICES_PER_TILE", "2")
    levels = {}
    for band in ("1", "0"):
        monkeypatch.setenv("RUV_BATCH_BAND", band)
        bt = lb.device_batch(precond=Precond.AMG, mixed=mixed)
        assert (bt.band_info()["band_charts"] > 0) == (band == "1")
        nl = bt.precond_info()["levels"]
        levels[band] = [bt.amg_level(l) for l in range(nl)]
        bt.close()
    assert len(levels["1"]) == len(levels["0"])
    for l, (a, b) in enumerate(zip(levels["1"], levels["0"])):
        np.testing.assert_array_equal(a["chart"], b["chart"])
        assert a["n_next"] == b["n_next"]
        if a["agg"] is not None:
            np.testing.assert_array_equal(a["agg"], b["agg"])
        if l > 0:
            assert (a["op"] != b["op"]).nnz == 0
            np.testing.assert_array_equal(a["op"].indptr, b["op"].indptr)
            np.testing.assert_array_equal(a["op"].indices, b["op"].indices)
            np.testing.assert_array_equal(a["op"].data, b["op"].data)


def test_batch_of_small_spd_systems_matches_golden():
    """One right-hand side per row (rhs_per_row = 1): the golden small SPD systems as one batch."""
    import os

    import scipy.sparse as sp

    from researchuv_amd import batch as B

    say("batch of the golden small SPD systems, R = 1")
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "spd_small.npz"))
    mats, bs, xs, off = [], [], [], [0]
    for k in range(int(z["count"])):
        v, c, p = z[f"m{k}_vals"], z[f"m{k}_cols"], z[f"m{k}_row_ptr"]
        m = len(p) - 1
        mats.append(sp.csr_matrix((v, c, p), shape=(m, m)))
        bs.append(z[f"m{k}_b"])
        xs.append(z[f"m{k}_x"])
        off.append(off[-1] + m)
    A = sp.block_diag(mats, format="csr")
    A.sort_indices()
    for mixed in (False, True):
        bt = B.Batch.from_csr(A.indptr.astype(np.int64), A.indices, A.data, off, mixed=mixed, R=1)
        x, info, it, res = bt.solve(np.concatenate(bs), max_iter=1000, tol=1e-12)
        for k in range(len(xs)):
            xk = x[off[k]:off[k + 1]]
            assert np.abs(xk - xs[k]).max() <= 1e-9 * max(1.0, np.abs(xs[k]).max()), (mixed, k)


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_amg_matches_per_chart_lu(orc, mixed):
    """The device-built aggregation AMG as the batch preconditioner: every chart against the oracle's LU, and
    far fewer iterations than Jacobi on the same batch."""
    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    say(f"batch AMG of 24 uneven charts mixed={mixed}")
    meshes = _meshes(24)
    lb = B.LscmBatchBuilder(meshes, angle_min_deg=180.0, threads=4)
    a = lb.arrays()
    tol = 1e-10 if not mixed else 1e-9
    bt = lb.device_batch(precond=Precond.AMG, mixed=mixed)
    info_pc = bt.precond_info()
    say(f"  hierarchy {info_pc}")
    assert info_pc["levels"] >= 2
    x, info, it, res = bt.solve(a["rhs"], max_iter=20000, tol=tol)
    ref = _oracle_solutions(orc, meshes)
    off = a["chart_row_off"]
    worst = 0.0
    for c in range(lb.nchart):
        xc = x[2 * off[c]:2 * off[c + 1]].reshape(-1, 2)
        rc = ref[c].reshape(-1, 2)
        worst = max(worst, np.linalg.norm(xc - rc, axis=1).max() / max(np.linalg.norm(rc, axis=1).max(), 1e-30))
    _, info_j, it_j, _ = lb.device_batch(precond=Precond.JACOBI, mixed=mixed).solve(a["rhs"], max_iter=20000, tol=tol)
    say(f"  worst chart uv err {worst:.2e}; max iters amg {info.iters} vs jacobi {info_j.iters}")
    assert worst < 1e-6
    if mixed:
        assert (res <= 1e-9).all()
    assert info.iters * 2 < info_j.iters


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_amg_coarsest_inverse_on_device(mixed):
    """The coarsest level's blocks are factored (Cholesky) and inverted on the device in FP64, one workgroup per
    chart: every chart's inverse against numpy's LU inverse of the same dense block, and no chart left on its
    diagonal (amg.rs:685-773 and precond.rs:439-498 factor the coarsest level exactly)."""
    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    meshes = _meshes(24, seed=6)
    lb = B.LscmBatchBuilder(meshes, angle_min_deg=180.0, threads=4)
    bt = lb.device_batch(precond=Precond.AMG, mixed=mixed)
    pc = bt.precond_info()
    say(f"coarsest blocks mixed={mixed}: {pc['level_rows'][-1]} rows over {lb.nchart} charts, {pc}")
    assert pc["coarsest_diagonal_charts"] == {"not_spd": 0, "coupled": 0, "too_large": 0}
    worst, sizes = 0.0, []
    for c in range(lb.nchart):
        blk, inv = bt.coarsest_block(c)
        if blk.size == 0:
            continue
        sizes.append(len(blk))
        np.testing.assert_array_equal(blk, blk.T)
        ref = np.linalg.inv(blk)
        worst = max(worst, np.abs(inv - ref).max() / np.abs(ref).max())
        assert np.abs(inv @ blk - np.eye(len(blk))).max() < 1e-10
    bt.close()
    say(f"  {len(sizes)} blocks of {min(sizes)}..{max(sizes)} rows, worst inverse rel diff vs numpy {worst:.2e}")
    assert max(sizes) >= 8
    assert worst < 1e-11


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_amg_device_hierarchy_invariants(mixed):
    """The device-built hierarchy level by level (its stored order): every row in exactly one aggregate, of its own
    chart, and every aggregate used; each coarse operator the Galerkin product P^T A P of the level above with the
    aggregate map as P (the tentative transfer of amg.rs:308-362, product of amg.rs:384-431); the entry sum kept;
    every level symmetric with a positive diagonal; the coarsest at most 64 rows a chart (the checks of
    amg.rs:799-942, applied to the device hierarchy).  Values as the cycle stores them: FP32-computed, held in
    FP64 (mixed=False) or bfloat16 (mixed=True), hence the tolerances."""
    import scipy.sparse as sp

    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    meshes = _meshes(24, seed=12)
    lb = B.LscmBatchBuilder(meshes, angle_min_deg=180.0, threads=4)
    a = lb.arrays(csr=True)
    n0 = lb.nrows
    bt = lb.device_batch(precond=Precond.AMG, mixed=mixed)
    nl = bt.precond_info()["levels"]
    say(f"device hierarchy mixed={mixed}: {bt.precond_info()['level_rows']}")
    assert nl >= 3
    rtol = 4e-3 if mixed else 2e-6
    op = sp.csr_matrix((a["vals"], a["cols"], a["row_ptr"]), shape=(n0, n0))
    prev = bt.amg_level(0)
    off = a["chart_row_off"]
    np.testing.assert_array_equal(prev["chart"], np.repeat(np.arange(lb.nchart), np.diff(off)))
    for lvl in range(1, nl):
        cur = bt.amg_level(lvl)
        agg, n = prev["agg"], cur["op"].shape[0]
        assert prev["n_next"] == n
        assert agg.min() >= 0 and agg.max() < n
        assert (np.bincount(agg, minlength=n) > 0).all()
        np.testing.assert_array_equal(cur["chart"][agg], prev["chart"])
        P = sp.csr_matrix((np.ones(len(agg)), (np.arange(len(agg)), agg)), shape=(len(agg), n))
        want = (P.T @ op @ P).tocsr()
        got = cur["op"]
        scale = abs(want).max()
        diff = abs(got - want).max()
        say(f"  level {lvl}: {n} rows, {got.nnz} entries, Galerkin max diff {diff / scale:.2e} of max |A|")
        assert diff <= rtol * scale
        assert abs(got.sum() - op.sum()) <= rtol * abs(op).sum()
        assert (got.diagonal() > 0).all()
        assert abs(got - got.T).max() <= rtol * scale
        op, prev = got, cur
    cnt = np.bincount(prev["chart"], minlength=lb.nchart)
    assert cnt.max() <= 64
    bt.close()


def test_batch_amg_general_slices_and_small_systems():
    """AMG over general (per-entry column) slices: the permuted chart of the wide-column test, and the golden
    small SPD systems with one right-hand side per row."""
    import os

    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from researchuv_amd import batch as B
    from researchuv_amd import lscm, meshgen
    from researchuv_amd.solver import Precond

    say("batch AMG on a permuted chart and on the small golden systems")
    pos, faces = meshgen.heightfield(120, 120, seed=4)
    s = lscm.segment(*lscm.weld(pos, faces, 1e-12), 180.0).system(0)
    n = s["nfree"]
    L = sp.csr_matrix((s["l_vals"], s["l_cols"], s["l_ptr"]), shape=(n, n))
    perm = np.random.default_rng(3).permutation(n)
    Lp = L[perm][:, perm].tocsr()
    Lp.sort_indices()
    rhs = s["rhs"].reshape(-1, 2)[perm]
    bt = B.Batch.from_csr(Lp.indptr.astype(np.int64), Lp.indices, Lp.data, [0, n], precond=Precond.AMG, mixed=True)
    x, info, it, res = bt.solve(rhs.reshape(-1), max_iter=20000, tol=1e-10)
    lu = spla.splu(Lp.tocsc())
    xr = np.stack([lu.solve(rhs[:, 0]), lu.solve(rhs[:, 1])], axis=1)
    err = np.linalg.norm(x.reshape(-1, 2) - xr, axis=1).max() / np.linalg.norm(xr, axis=1).max()
    say(f"  permuted chart: uv err {err:.2e}, iters {info.iters}, levels {bt.precond_info()['levels']}")
    assert err < 1e-6
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "spd_small.npz"))
    mats, bs, xs, off = [], [], [], [0]
    for k in range(int(z["count"])):
        v, c, p = z[f"m{k}_vals"], z[f"m{k}_cols"], z[f"m{k}_row_ptr"]
        mats.append(sp.csr_matrix((v, c, p), shape=(len(p) - 1,) * 2))
        bs.append(z[f"m{k}_b"])
        xs.append(z[f"m{k}_x"])
        off.append(off[-1] + len(p) - 1)
    A = sp.block_diag(mats, format="csr")
    A.sort_indices()
    for mixed in (False, True):
        bt = B.Batch.from_csr(A.indptr.astype(np.int64), A.indices, A.data, off, precond=Precond.AMG, mixed=mixed, R=1)
        x, info, it, res = bt.solve(np.concatenate(bs), max_iter=1000, tol=1e-12)
        for k in range(len(xs)):
            assert np.abs(x[off[k]:off[k + 1]] - xs[k]).max() <= 1e-9 * max(1.0, np.abs(xs[k]).max()), (mixed, k)


def test_stretched_grid_amg_against_superlu_and_ic0():
    """The ill-conditioned chart (column spacing growing 1 -> 100, 5% flipped diagonals) at 160 x 160: the
    batched AMG reaches the SuperLU solution with far fewer iterations than Jacobi; the plan path's IC(0)
    (the reference's preconditioner for this case) is timed beside it for the record."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from researchuv_amd import batch as B
    from researchuv_amd import lscm, meshgen, solver
    from researchuv_amd.solver import Precond

    say("stretched 160x160 grid: batched AMG vs SuperLU")
    pos, faces = meshgen.stretched_grid(160, stretch=100.0, flip_frac=0.05, seed=1)
    seg = lscm.segment(*lscm.weld(pos, faces, 1e-12), 180.0)
    assert len(seg) == 1
    s = seg.system(0)
    n = s["nfree"]
    L = sp.csr_matrix((s["l_vals"], s["l_cols"], s["l_ptr"]), shape=(n, n))
    r = s["rhs"].reshape(-1, 2)
    lu = spla.splu(L.tocsc())
    xr = np.stack([lu.solve(r[:, 0]), lu.solve(r[:, 1])], axis=1)
    its = {}
    for pc in (Precond.AMG, Precond.JACOBI):
        bt = B.Batch.from_csr(s["l_ptr"].astype(np.int64), s["l_cols"], s["l_vals"], [0, n], precond=pc, mixed=True)
        x, info, it, res = bt.solve(s["rhs"], max_iter=200000, tol=1e-10)
        err = np.linalg.norm(x.reshape(-1, 2) - xr, axis=1).max() / np.linalg.norm(xr, axis=1).max()
        its[pc.name] = int(info.iters)
        say(f"  {pc.name}: iters {info.iters}, true {info.true_rel_resid:.1e}, uv err {err:.2e}")
        assert err < 1e-6
    A = sp.kron(L, sp.identity(2), format="csr")  # the interleaved A = L (x) I2 the drop-in receives
    A.sort_indices()
    gs = solver.GpuSolver.new()
    out = gs.cg_solve_mixed(solver.CsrMatrix(A.data, A.indices.astype(np.int32), A.indptr.astype(np.int32)),
                            s["rhs"], 200000, 1e-10, Precond.IC0)
    assert out is not None, gs.last_error
    err = np.linalg.norm(out[0].reshape(-1, 2) - xr, axis=1).max() / np.linalg.norm(xr, axis=1).max()
    say(f"  plan IC0: iters {out[1].iters}, uv err {err:.2e}")
    assert err < 1e-6
    assert its["AMG"] * 5 < its["JACOBI"]


def test_batch_amg_is_deterministic(monkeypatch):
    """The device-built hierarchy and the batched solve are bitwise reproducible: two solves with one batch,
    and two batches built from the same input (schedule held fixed), give identical solutions."""
    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    say("batch AMG determinism")
    monkeypatch.setenv("RUV_BATCH_SLICES_PER_TILE", "2")
    meshes = _meshes(12, seed=21)
    lb = B.LscmBatchBuilder(meshes, threads=2)
    a = lb.arrays()
    xs = []
    for k in range(2):
        bt = lb.device_batch(precond=Precond.AMG, mixed=True)
        x1, i1, _, _ = bt.solve(a["rhs"], max_iter=20000, tol=1e-10)
        bt.setup_precond()  # rebuild in place
        x2, i2, _, _ = bt.solve(a["rhs"], max_iter=20000, tol=1e-10)
        np.testing.assert_array_equal(x1, x2)
        assert bt.precond_info()["level_rows"] == bt.precond_info()["level_rows"]
        xs.append(x1)
        bt.close()
    np.testing.assert_array_equal(xs[0], xs[1])


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_step_profile(mixed):
    """The whole-batch step profile marks every kernel of the preconditioned step once per level and step, with
    positive times and the byte counts the level sizes give; a solve after it (state overwritten) is bitwise the
    solve before it."""
    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    say(f"batch step profile mixed={mixed}")
    meshes = _meshes(12, seed=23)
    lb = B.LscmBatchBuilder(meshes, threads=2)
    a = lb.arrays()
    bt = lb.device_batch(precond=Precond.AMG, mixed=mixed)
    x1, _, _, _ = bt.solve(a["rhs"], max_iter=20000, tol=1e-10)
    prof = bt.step_profile(3)
    rows = bt.precond_info()["level_rows"]
    nl = len(rows)
    kinds = [e["kernel"] for e in prof]
    for k in ("kb_dir", "kb_spmv", "kb_update", "kc_resid0", "kc_restrict", "kc_coarsest", "kc_prolong0", "kc_post0"):
        assert kinds.count(k) == 1, (k, kinds)
    assert sorted(e["level"] for e in prof if e["kernel"] == "kc_down") == list(range(1, nl - 1))
    assert sorted(e["level"] for e in prof if e["kernel"] == "kc_post") == list(range(1, nl - 1))
    assert all(e["seconds"] > 0 and e["bytes"] > 0 for e in prof)
    # stored bytes cover the padding slots; the finest-operator passes of band charts store each mirrored pair once
    finest = ("kb_spmv", "kc_resid0", "kc_post0")
    assert bt.band_info()["band_charts"] > 0
    assert all(e["stored_bytes"] >= e["bytes"] for e in prof if e["kernel"] not in finest)
    assert all(0 < e["stored_bytes"] < e["bytes"] for e in prof if e["kernel"] in finest)
    vb = 2 * (4 if mixed else 8)
    n = rows[0]
    upd = [e for e in prof if e["kernel"] == "kb_update"][0]
    assert upd["bytes"] == 3 * n * vb  # q, r read; r written
    dr = [e for e in prof if e["kernel"] == "kb_dir"][0]
    assert dr["bytes"] == 5 * n * vb  # z, old p, x read; p, x written (x += alpha p of the previous step)
    coarsest = [e for e in prof if e["kernel"] == "kc_coarsest"][0]
    assert coarsest["level"] == nl - 1
    x2, _, _, _ = bt.solve(a["rhs"], max_iter=20000, tol=1e-10)
    np.testing.assert_array_equal(x1, x2)
    bt.close()


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_amg_degenerate_charts(mixed):
    """The device AMG over charts of 0, 1, 2 and 3 rows, a permuted chart (32-bit columns on every level) and a
    trailing empty chart: every chart against SuperLU."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from researchuv_amd import batch as B
    from researchuv_amd import lscm, meshgen
    from researchuv_amd.solver import Precond

    say(f"AMG with degenerate charts mixed={mixed}")
    pos, faces = meshgen.heightfield(120, 120, seed=8)
    s = lscm.segment(*lscm.weld(pos, faces, 1e-12), 180.0).system(0)
    n = s["nfree"]
    L = sp.csr_matrix((s["l_vals"], s["l_cols"], s["l_ptr"]), shape=(n, n))
    perm = np.random.default_rng(4).permutation(n)
    Lp = L[perm][:, perm].tocsr()
    rhs_p = s["rhs"].reshape(-1, 2)[perm]
    blocks = [sp.csr_matrix((0, 0)), sp.csr_matrix([[4.0]]), sp.csr_matrix([[3.0, -1.0], [-1.0, 2.0]]),
              sp.csr_matrix([[2.0, -1.0, 0.0], [-1.0, 2.0, -1.0], [0.0, -1.0, 2.0]]), Lp, sp.csr_matrix((0, 0))]
    rng = np.random.default_rng(7)
    rhs = [rng.normal(size=(b.shape[0], 2)) for b in blocks]
    rhs[4] = rhs_p
    A = sp.block_diag([b for b in blocks if b.shape[0]], format="csr")
    A.sort_indices()
    off = np.concatenate([[0], np.cumsum([b.shape[0] for b in blocks])])
    b = np.concatenate([r.reshape(-1) for r in rhs])
    bt = B.Batch.from_csr(A.indptr.astype(np.int64), A.indices, A.data, off, precond=Precond.AMG, mixed=mixed)
    assert bt.precond_info()["levels"] >= 2
    x, info, it, res = bt.solve(b, max_iter=20000, tol=1e-10)
    for c, blk in enumerate(blocks):
        m = blk.shape[0]
        if m == 0:
            assert it[c] == 0
            continue
        lu = spla.splu(blk.tocsc())
        xr = np.stack([lu.solve(rhs[c][:, 0]), lu.solve(rhs[c][:, 1])], axis=1)
        xc = x[2 * off[c]:2 * off[c + 1]].reshape(-1, 2)
        err = np.linalg.norm(xc - xr, axis=1).max() / np.linalg.norm(xr, axis=1).max()
        say(f"  chart {c} ({m} rows): uv err {err:.2e}, iters {it[c]}")
        assert err < 1e-6, c


def test_batch_refinement_rounds_and_live_lists(orc, monkeypatch):
    """Refinement rounds that go on from the previous round's direction reach the same solutions as fresh
    restarts; with the continuation test never passing (gap 0, beta = 0) they are bitwise the restarts.  Listing
    only the running charts' slices for the coarse launches, and folding the coarse prolongation into the
    post-smoothing pass, and forming the chart-wide sums once per chart, change no arithmetic: bitwise the full
    launches, the separate passes and the per-workgroup sums."""
    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    say("batch AMG refinement: continued / restarted rounds, listed / full coarse launches")
    meshes = _meshes(24, seed=5)
    lb = B.LscmBatchBuilder(meshes, angle_min_deg=180.0, threads=4)
    a = lb.arrays()
    tol = 1e-9

    def solve(**env):
        for k in ("RUV_BATCH_RESTART", "RUV_BATCH_CONT_GAP", "RUV_BATCH_LIST_BELOW", "RUV_BATCH_AMG_FUSE_ROWS",
                  "RUV_BATCH_BIG_TILES"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        bt = lb.device_batch(precond=Precond.AMG, mixed=True)
        return bt.solve(a["rhs"], max_iter=20000, tol=tol)

    x_c, i_c, it_c, r_c = solve()
    x_r, i_r, it_r, r_r = solve(RUV_BATCH_RESTART="1")
    x_g, i_g, it_g, r_g = solve(RUV_BATCH_CONT_GAP="0")
    x_n, _, it_n, _ = solve(RUV_BATCH_LIST_BELOW="0")
    x_a, _, it_a, _ = solve(RUV_BATCH_LIST_BELOW="2")
    x_u, _, it_u, _ = solve(RUV_BATCH_AMG_FUSE_ROWS="0")  # prolongation always a pass of its own
    x_v, _, it_v, _ = solve(RUV_BATCH_AMG_FUSE_ROWS="1000000000")  # always inside the post-smoothing
    x_s, _, it_s, _ = solve(RUV_BATCH_BIG_TILES="1")  # chart-wide sums once per chart (kb_chart_sums)
    say(f"  max iters continued {i_c.iters} restarted {i_r.iters}; worst resid {r_c.max():.2e} / {r_r.max():.2e}")
    assert (r_c <= tol).all() and (r_r <= tol).all()
    assert np.array_equal(x_g, x_r) and np.array_equal(it_g, it_r)
    assert np.array_equal(x_n, x_a) and np.array_equal(it_n, it_a)
    assert np.array_equal(x_a, x_c) and np.array_equal(it_a, it_c)
    assert np.array_equal(x_u, x_c) and np.array_equal(it_u, it_c)
    assert np.array_equal(x_v, x_c) and np.array_equal(it_v, it_c)
    assert np.array_equal(x_s, x_c) and np.array_equal(it_s, it_c)
    ref = _oracle_solutions(orc, meshes)
    off = a["chart_row_off"]
    for x in (x_c, x_r):
        worst = 0.0
        for c in range(lb.nchart):
            xc = x[2 * off[c]:2 * off[c + 1]].reshape(-1, 2)
            rc = ref[c].reshape(-1, 2)
            worst = max(worst, np.linalg.norm(xc - rc, axis=1).max() / max(np.linalg.norm(rc, axis=1).max(), 1e-30))
        assert worst < 1e-6


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_reports_a_chart_whose_step_breaks_down(mixed):
    """A chart whose first step meets p.Ap = 0 (the indefinite block diag(1, -1), no preconditioner) stops as broken
    down in the update pass while the SPD chart beside it runs on; the batch call reports the breakdown
    (solver.rs:520-526 stops on a zero p.Ap)."""
    import scipy.sparse as sp

    from researchuv_amd import _lib
    from researchuv_amd import batch as B
    from researchuv_amd.solver import Precond

    say(f"batch breakdown in the update pass mixed={mixed}")
    spd = sp.diags([-np.ones(63), 2.0 * np.ones(64), -np.ones(63)], [-1, 0, 1], format="csr")
    bad = sp.csr_matrix(np.diag([1.0, -1.0]))
    A = sp.block_diag([spd, bad], format="csr")
    A.sort_indices()
    off = [0, 64, 66]
    bt = B.Batch.from_csr(A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, off, precond=Precond.NONE,
                          mixed=mixed, R=1)
    rhs = np.ones(66)
    with pytest.raises(_lib.RuvError) as e:
        bt.solve(rhs, max_iter=1000, tol=1e-10)
    assert "BREAKDOWN" in str(e.value) or "broke down" in str(e.value)
    bt.close()
    # the SPD chart alone solves
    A1 = spd.tocsr()
    A1.sort_indices()
    b1 = B.Batch.from_csr(A1.indptr.astype(np.int64), A1.indices.astype(np.int32), A1.data, [0, 64],
                          precond=Precond.NONE, mixed=mixed, R=1)
    x, info, it, res = b1.solve(np.ones(64), max_iter=1000, tol=1e-10)
    assert np.abs(spd @ x - 1.0).max() < 1e-8
    b1.close()
