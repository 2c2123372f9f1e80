"""The batched unfold stage (pipeline.rs:171-194 for many meshes in one call) against the oracle's worker-thread
path (orc_par_unfold, the CPU sparse LU per chart) and its per-chart metrics: chart and vertex indexing bit for
bit, uv within the reference's 1e-6 (lscm.rs:749-758), the UnfoldResult fields, rectangularize and
chart_distortion (whose fold count is recut::chart_folds', recut.rs:71-80)."""
import numpy as np
import pytest

from tests import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _hf(n, seed):
    from researchuv_amd import meshgen

    return meshgen.heightfield(n, n, seed=seed)


GROUPS = {
    30.0: lambda orc: [orc.cube(6), orc.grid(32, 32), meshes.torus_annulus(2.0, 0.7, 24, 16), _hf(40, 8)],
    180.0: lambda orc: [meshes.uv_sphere(16, 8), _hf(48, 3), _hf(36, 4)],
}


@pytest.mark.parametrize("angle", list(GROUPS))
def test_unfold_batch_matches_oracle_par_unfold(orc, angle):
    from researchuv_amd import lscm
    from researchuv_amd.solver import Precond

    ms = GROUPS[angle](orc)
    opts = lscm.UnfoldOptions(precond=Precond.AMG, mixed=True, tol=1e-10)
    res, info = lscm.unfold_batch(ms, opts, angle_min_deg=angle, threads=4)
    say(f"angle {angle}: {len(ms)} meshes -> {len(res)} charts, max iters {info.iters}, "
        f"worst true resid {info.true_rel_resid:.2e}")
    c = 0
    worst = 0.0
    for m, (pos, faces) in enumerate(ms):
        pw, fw = orc.weld(pos, faces, 1e-12)
        co = orc.Charts(pw, fw, angle)
        uvs = co.par_unfold(3)
        for i in range(len(co)):
            r = res[c]
            ch = co.chart(i)
            assert r.mesh == m
            np.testing.assert_array_equal(r.vertex_ids, ch["vertex_ids"])
            uv_cpu, st = co.unfold(i)
            np.testing.assert_array_equal(uv_cpu, uvs[i])
            err = np.linalg.norm(r.uv - uv_cpu, axis=1).max() / np.linalg.norm(uv_cpu, axis=1).max()
            worst = max(worst, err)
            assert err < 1e-6, (m, i, err)
            assert r.iters == st["iters"] == 2
            assert r.area3d == st["area3d"] and r.target == st["target"]
            assert abs(r.err_accum - st["err_accum"]) <= 1e-6 * st["err_accum"]
            # the stage's metrics are the oracle's metrics of the stage's own uv (the fold count of a chart whose
            # signed areas nearly cancel -- the annulus -- flips with any rounding of the uv, so it is compared on
            # the same uv): rectified uv, extents and counts bit for bit, the means to the rounding of their sums
            p3 = pw[ch["vertex_ids"]]
            met = orc.chart_distortion(p3, ch["tris"], r.uv)
            for k in ("conformal_mean", "conformal_max", "area_ratio_mean"):
                assert abs(getattr(r, k) - met[k]) <= 1e-12 * max(abs(met[k]), 1.0), (m, i, k)
            assert r.flips == met["flips"] and r.folds == met["folds"]
            bnd = np.searchsorted(ch["vertex_ids"], ch["border"])
            q, ext = orc.rectangularize(r.uv, bnd)
            np.testing.assert_array_equal(r.rectified, q)
            np.testing.assert_array_equal(r.ext, ext)
            c += 1
    assert c == len(res)
    say(f"  worst chart uv err {worst:.2e}")


def test_unfold_batch_is_independent_of_worker_count(orc):
    """Results in input order, identical for any worker count (pipeline.rs:369-385)."""
    from researchuv_amd import lscm
    from researchuv_amd.solver import Precond

    ms = GROUPS[30.0](orc)
    opts = lscm.UnfoldOptions(precond=Precond.AMG, mixed=True, tol=1e-10)
    a, _ = lscm.unfold_batch(ms, opts, threads=1)
    b, _ = lscm.unfold_batch(ms, opts, threads=5)
    assert len(a) == len(b)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.uv, y.uv)
        np.testing.assert_array_equal(x.rectified, y.rectified)
        assert (x.iters, x.err_accum, x.folds, x.conformal_mean) == (y.iters, y.err_accum, y.folds, y.conformal_mean)
