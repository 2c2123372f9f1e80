"""GPU tests of the product's own per-chart path: device assembly (bitwise the host assembly, which is
bitwise the oracle's) and the unfold driver against the oracle's CPU sparse-LU unfold (lscm.rs:733-760)."""
import numpy as np
import pytest

from tests import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


CASES = {
    "cube6": (lambda orc: orc.cube(6), 30.0),
    "grid32": (lambda orc: orc.grid(32, 32), 30.0),
    "annulus": (lambda orc: meshes.torus_annulus(2.0, 0.7, 24, 16), 30.0),
    "sphere": (lambda orc: meshes.uv_sphere(16, 8), 180.0),
}


def _heightfield(n, seed=3):
    from researchuv_amd import meshgen

    return meshgen.heightfield(n, n, seed=seed)


@pytest.mark.parametrize("name", list(CASES) + ["heightfield48"])
def test_device_assembly_is_bitwise_the_host_assembly(orc, name):
    from researchuv_amd import lscm

    say(f"device assembly {name}")
    if name == "heightfield48":
        pos, faces = _heightfield(48)
        angle = 180.0
    else:
        fn, angle = CASES[name]
        pos, faces = fn(orc)
    pw, fw = lscm.weld(pos, faces, 1e-12)
    seg = lscm.segment(pw, fw, angle)
    for i in range(len(seg)):
        h = seg.system(i)
        d = seg.system_device(i)
        for k in ("l_vals", "l_cols", "l_ptr", "rhs"):
            assert np.array_equal(h[k], d[k]), (name, i, k)
        assert h["edges"] == d["edges"]


def uv_rel_err(a, b):
    return np.linalg.norm(a - b, axis=1).max() / max(np.linalg.norm(b, axis=1).max(), 1e-30)


@pytest.mark.parametrize("name", list(CASES) + ["heightfield40"])
@pytest.mark.parametrize("mixed", [False, True])
def test_unfold_chart_gpu_matches_cpu_lu(orc, name, mixed):
    """The reference's acceptance test (lscm.rs:733-760) through the product's own pipeline: weld, segment,
    pins, device assembly, device PCG, driver loop; compared with the oracle's SolverBackend::Cpu unfold."""
    from researchuv_amd import lscm

    say(f"unfold {name} mixed={mixed}")
    if name == "heightfield40":
        pos, faces = _heightfield(40, seed=8)
        angle = 180.0
    else:
        fn, angle = CASES[name]
        pos, faces = fn(orc)
    pw, fw = lscm.weld(pos, faces, 1e-12)
    seg = lscm.segment(pw, fw, angle)
    pwo, fwo = orc.weld(pos, faces, 1e-12)
    co = orc.Charts(pwo, fwo, angle)
    assert len(seg) == len(co)
    opts = lscm.UnfoldOptions(min_vars=0, mixed=mixed, tol=1e-12 if not mixed else 1e-10)
    for i in range(len(seg)):
        uv, st, info = seg.unfold_chart(i, opts)
        uv_cpu, st_cpu = co.unfold(i)
        err = uv_rel_err(uv, uv_cpu)
        assert st["iters"] == st_cpu["iters"] == 2
        assert err < 1e-6, (name, i, err)


def test_unfold_respects_the_size_gate(orc):
    from researchuv_amd import RuvError, lscm

    pw, fw = lscm.weld(*orc.cube(6), 1e-12)
    seg = lscm.segment(pw, fw, 30.0)
    with pytest.raises(RuvError) as e:
        seg.unfold_chart(0, lscm.UnfoldOptions())  # 50 unknowns < GPU_MIN_VARS = 1024 (lscm.rs:462,479)
    assert e.value.status == -2
