"""Bit-exact parity of the product's host side (charting, pins, weights, assembly) with the oracle.

Chart and vertex indexing and pin choice must match the reference bit for bit; the assembled free block
A = L (x) I2 and its right-hand side are compared bitwise too.  Runs on the CPU (no device calls)."""
import numpy as np
import pytest

from tests import meshes


def _heightfield(nx, ny, seed=7, amp=0.1, jitter=0.2):
    from researchuv_amd import meshgen

    return meshgen.heightfield(nx, ny, seed=seed, amp=amp, jitter=jitter)


CASES = {
    "cube6": lambda orc: orc.cube(6),
    "grid32": lambda orc: orc.grid(32, 32),
    "grid7x3": lambda orc: orc.grid(7, 3),
    "cylinder": lambda orc: meshes.cylinder(1.0, 2.0, 32, 16),
    "annulus": lambda orc: meshes.torus_annulus(2.0, 0.7, 24, 16),
    "sphere": lambda orc: meshes.uv_sphere(16, 8),
    "torus": lambda orc: meshes.torus(2.0, 0.7, 16, 10),
    "pinched": lambda orc: meshes.pinched_grid(),
    "pinched_tie": lambda orc: meshes.pinched_grid_tie(),
}
ANGLES = {"sphere": 180.0, "torus": 180.0}


@pytest.fixture(scope="module")
def ra_lscm():
    from researchuv_amd import lscm

    return lscm


def _both(orc, ra_lscm, pos, faces, angle):
    pw_o, fw_o = orc.weld(pos, faces, 1e-12)
    pw_p, fw_p = ra_lscm.weld(pos, faces, 1e-12)
    assert np.array_equal(pw_o, pw_p) and np.array_equal(fw_o, fw_p)
    return orc.Charts(pw_o, fw_o, angle), ra_lscm.segment(pw_p, fw_p, angle)


@pytest.mark.parametrize("name", list(CASES))
def test_indexing_pins_and_assembly_are_bit_exact(orc, ra_lscm, name):
    pos, faces = CASES[name](orc)
    co, cp = _both(orc, ra_lscm, pos, faces, ANGLES.get(name, 30.0))
    assert len(co) == len(cp)
    assert np.array_equal(co.cut(), np.zeros((0, 2), np.uint32)) or cp.n_cut() == len(co.cut())
    for i in range(len(co)):
        a, b = co.chart(i), cp.chart(i)
        for k in ("face_ids", "vertex_ids", "border"):
            assert np.array_equal(a[k], getattr(b, k)), k
        assert np.array_equal(a["tris"], b.tris)
        assert len(a["border_loops"]) == len(b.border_loops)
        for x, y in zip(a["border_loops"], b.border_loops):
            assert np.array_equal(x, y)
        assert a["area3d"] == b.area3d
        so, sp_ = co.system(i), cp.system(i)
        assert np.array_equal(so["p2"], sp_["p2"])
        assert np.array_equal(so["pin_ids"], sp_["pin_ids"])
        assert np.array_equal(so["pin_uv"], sp_["pin_uv"])
        assert np.array_equal(so["free_ids"], sp_["free_ids"])
        # A = L (x) I2: the oracle's interleaved CSR, row 2p, is row p of L on columns 2q
        n = 2 * sp_["nfree"]
        ev = np.arange(0, n, 2)
        lo, hi = so["row_ptr"][ev], so["row_ptr"][ev + 1]
        cols = np.concatenate([so["cols"][a_:b_] for a_, b_ in zip(lo, hi)]) if n else np.zeros(0, np.int32)
        vals = np.concatenate([so["vals"][a_:b_] for a_, b_ in zip(lo, hi)]) if n else np.zeros(0)
        assert np.array_equal(cols // 2, sp_["l_cols"]) and np.array_equal(vals, sp_["l_vals"])
        assert np.array_equal(np.diff(sp_["l_ptr"]), hi - lo)
        # odd rows repeat the even ones on odd columns
        lo1, hi1 = so["row_ptr"][ev + 1], so["row_ptr"][ev + 2]
        vals1 = np.concatenate([so["vals"][a_:b_] for a_, b_ in zip(lo1, hi1)]) if n else np.zeros(0)
        assert np.array_equal(vals1, vals)
        assert np.array_equal(so["rhs"][:n], sp_["rhs"])


def test_heightfield_chart_is_bit_exact(orc, ra_lscm):
    pos, faces = _heightfield(24, 20)
    co, cp = _both(orc, ra_lscm, pos, faces, 180.0)
    assert len(co) == len(cp) == 1
    so, sp_ = co.system(0), cp.system(0)
    for k in ("p2", "pin_ids", "pin_uv", "free_ids"):
        assert np.array_equal(so[k], sp_[k]), k
    n = 2 * sp_["nfree"]
    assert np.array_equal(so["rhs"][:n], sp_["rhs"])
    assert len(so["vals"]) == 2 * len(sp_["l_vals"])


def test_batch_builder_matches_per_chart_systems_and_thread_counts(ra_lscm):
    """The worker-thread batch builder (exec::par_map semantics: results in input order, identical for any
    worker count) lays out exactly the per-chart systems."""
    from researchuv_amd import batch as B
    from researchuv_amd import meshgen

    sizes = meshgen.patch_sizes(12, vmin=64, vmax=900, seed=4)
    meshes = [meshgen.heightfield(int(a), int(b), seed=40 + k) for k, (a, b) in enumerate(sizes)]
    a1 = B.LscmBatchBuilder(meshes, threads=1).arrays(csr=True)
    a4 = B.LscmBatchBuilder(meshes, threads=4).arrays(csr=True)
    for k in a1:
        assert np.array_equal(a1[k], a4[k]), k
    off = a1["chart_row_off"]
    c = 0
    for pos, faces in meshes:
        pw, fw = ra_lscm.weld(pos, faces, 1e-12)
        seg = ra_lscm.segment(pw, fw, 180.0)
        for i in range(len(seg)):
            s = seg.system(i)
            r0, r1 = off[c], off[c + 1]
            assert r1 - r0 == s["nfree"]
            assert np.array_equal(a1["rhs"][2 * r0:2 * r1], s["rhs"])
            lo, hi = a1["row_ptr"][r0], a1["row_ptr"][r1]
            assert np.array_equal(a1["vals"][lo:hi], s["l_vals"])
            assert np.array_equal(a1["cols"][lo:hi] - r0, s["l_cols"])
            c += 1
    assert c == len(off) - 1
