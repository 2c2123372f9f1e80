"""The allocation-free chart metrics entry point (ruv_chart_metrics_dev_noalloc) against ruv_chart_metrics_dev:
the same kernels on a fixed device scratch array, so every metric and count must be bitwise equal.  Two calls in
a row, the second on a smaller input (the scratch keeps the first call's contents), then an input beyond the
scratch, which the host rejects before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _charts(count, seed, vmax):
    from researchuv_amd import lscm, meshgen

    out = []
    for pos, faces in meshgen.patch_batch(count, seed=seed, vmax=vmax):
        pw, fw = lscm.weld(pos, faces, 1e-12)
        seg = lscm.segment(pw, fw, 180.0)
        for i in range(len(seg)):
            ch = seg.chart(i)
            vids = np.asarray(ch.vertex_ids, np.int64)
            p3 = pw[vids]
            uv = np.ascontiguousarray(p3[:, :2] * np.array([1.0, 1.3]))  # an affine stand-in for an unfold
            out.append((p3, np.asarray(ch.tris, np.uint32), uv))
    return out


def _device(charts):
    import torch

    dev = torch.device("cuda:0")
    off = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.int64)  # noqa: E731
    v_off, t_off = off([c[0] for c in charts]), off([c[1] for c in charts])
    d = {k: torch.from_numpy(np.ascontiguousarray(a)).to(dev) for k, a in [
        ("v_off", v_off), ("t_off", t_off), ("p3", np.concatenate([c[0] for c in charts])),
        ("uv", np.concatenate([c[2] for c in charts])),
        ("tris", np.concatenate([c[1] for c in charts]).view(np.int32))]}
    torch.cuda.synchronize()
    return d


def _call(fn, n, d):
    import torch

    from researchuv_amd import _lib

    met = torch.full((n, 3), -1.0, dtype=torch.float64, device="cuda:0")
    cnt = torch.full((n, 2), -1, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    V = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    _lib.check(fn(n, V(d["v_off"]), V(d["p3"]), V(d["uv"]), V(d["t_off"]), V(d["tris"]), V(met), V(cnt)))
    return met.cpu().numpy(), cnt.cpu().numpy()


def test_noalloc_matches_dev_on_two_calls():
    from researchuv_amd import metrics as M

    L = M.lib()
    for k, (count, seed, vmax) in enumerate([(24, 21, 20000), (6, 22, 3000)]):
        charts = _charts(count, seed, vmax)
        d = _device(charts)
        n = len(charts)
        say(f"call {k}: {n} charts, {int(d['t_off'][-1])} triangles")
        m_ref, c_ref = _call(L.ruv_chart_metrics_dev, n, d)
        m_na, c_na = _call(L.ruv_chart_metrics_dev_noalloc, n, d)
        assert (c_ref >= 0).all() and np.isfinite(m_ref).all()
        np.testing.assert_array_equal(m_na, m_ref)
        np.testing.assert_array_equal(c_na, c_ref)


def test_noalloc_rejects_input_beyond_its_scratch():
    import torch

    from researchuv_amd import _lib
    from researchuv_amd import metrics as M

    dev = torch.device("cuda:0")
    # one chart claiming 2e8 triangles: about 195 000 tiles, more than the fixed scratch holds; the host reads
    # the count back and refuses before launching anything, so the arrays it would read are never touched
    v_off = torch.tensor([0, 3], dtype=torch.int64, device=dev)
    t_off = torch.tensor([0, 200_000_000], dtype=torch.int64, device=dev)
    p3 = torch.zeros((3, 3), dtype=torch.float64, device=dev)
    uv = torch.zeros((3, 2), dtype=torch.float64, device=dev)
    tris = torch.zeros(3, dtype=torch.int32, device=dev)
    met = torch.zeros((1, 3), dtype=torch.float64, device=dev)
    cnt = torch.zeros((1, 2), dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    V = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    rc = M.lib().ruv_chart_metrics_dev_noalloc(1, V(v_off), V(p3), V(uv), V(t_off), V(tris), V(met), V(cnt))
    assert rc == -2, rc
    assert "scratch" in _lib.last_error()
