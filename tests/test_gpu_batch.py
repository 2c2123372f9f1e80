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


def test_batch_is_independent_of_composition(orc):
    """Each chart runs its own PCG: solving a chart alone or inside a batch gives the same iterates."""
    from researchuv_amd import batch as B

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
