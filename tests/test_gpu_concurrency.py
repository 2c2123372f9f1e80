"""Concurrent callers.  The reference called the GPU solver from exec::par_map workers (pipeline.rs:192-194)
under `unsafe impl Send + Sync` (solver.rs:77-80); the C ABI promises that every entry point is thread-safe and
that each call runs on its own stream.  Four threads call the per-chart solve (three preconditioners), the
batched solve (one batch handle per thread) and the device-pointer chart metrics at once, three times over;
every result must equal the serial call's bit for bit."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _systems():
    import scipy.sparse as sp

    from researchuv_amd import lscm, meshgen, solver

    out = []
    for k, (nx, ny) in enumerate([(40, 40), (52, 36), (64, 30), (30, 72)]):
        pos, faces = meshgen.heightfield(nx, ny, seed=400 + k)
        pw, fw = lscm.weld(pos, faces, 1e-12)
        seg = lscm.segment(pw, fw, 180.0)
        ch = seg.chart(0)
        s = seg.system(0)
        vids = np.asarray(ch.vertex_ids, np.int64)
        n = s["nfree"]
        L = sp.csr_matrix((s["l_vals"], s["l_cols"], s["l_ptr"]), shape=(n, n))
        A = sp.kron(L, sp.identity(2, format="csr"), format="csr")  # the interleaved free block (lscm.rs:616-629)
        A.sort_indices()
        out.append(dict(sys=s, A=solver.CsrMatrix(A.data, A.indices, A.indptr), p3=pw[vids],
                        tris=np.asarray(ch.tris, np.uint32), mesh=(pos, faces)))
    return out


def _work(k, item, precs):
    """Every call of one worker: per-chart solves, a batched solve, device metrics.  Returns the results."""
    import ctypes as C

    import torch

    from researchuv_amd import _lib, batch as B, solver
    from researchuv_amd import metrics as M

    s = item["sys"]
    a = item["A"]
    n = a.rows()
    gs = solver.GpuSolver.new()
    res = {}
    for pc in precs:
        r = gs.cg_solve_precond(a, s["rhs"][:n], 16384, 1e-12, pc)  # tol of lscm.rs:501
        assert r is not None, gs.last_error
        res[f"cg_{pc.name}"] = r[0]
    lb = B.LscmBatchBuilder([item["mesh"]], angle_min_deg=180.0, threads=1)
    ar = lb.arrays()
    bt = lb.device_batch(precond=solver.Precond.AMG, mixed=True)
    x, info, it, rr = bt.solve(ar["rhs"], max_iter=20000, tol=1e-10)
    bt.close()
    lb.close()
    res["batch"] = x
    # device-pointer metrics: inputs written (and complete) on the null stream before the call
    dev = torch.device("cuda:0")
    uv = np.ascontiguousarray(item["p3"][:, :2]) * (1.0 + 0.01 * k)
    d = {name: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for name, v in [
        ("v_off", np.array([0, len(uv)], np.int64)), ("t_off", np.array([0, len(item["tris"])], np.int64)),
        ("p3", item["p3"]), ("uv", uv), ("tris", item["tris"].view(np.int32))]}
    torch.cuda.synchronize()
    met = torch.zeros((1, 3), dtype=torch.float64, device=dev)
    cnt = torch.zeros((1, 2), dtype=torch.int64, device=dev)
    V = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    _lib.check(M.lib().ruv_chart_metrics_dev(1, V(d["v_off"]), V(d["p3"]), V(d["uv"]), V(d["t_off"]), V(d["tris"]),
                                             V(met), V(cnt)))
    res["metrics"] = np.concatenate([met.cpu().numpy().ravel(), cnt.cpu().numpy().ravel().astype(np.float64)])
    return res


def test_concurrent_callers_bitwise_equal_serial():
    from researchuv_amd.solver import Precond

    precs = [Precond.JACOBI, Precond.IC0, Precond.AMG]
    items = _systems()
    say("serial calls")
    serial = [_work(k, it, precs) for k, it in enumerate(items)]
    for rep in range(3):
        say(f"4 threads at once, round {rep}")
        out = [None] * len(items)
        errs = []
        barrier = threading.Barrier(len(items))

        def run(k):
            try:
                barrier.wait()
                out[k] = _work(k, items[k], precs)
            except Exception as e:  # reported below
                errs.append(repr(e))

        th = [threading.Thread(target=run, args=(k,)) for k in range(len(items))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errs, errs
        for k in range(len(items)):
            for key, v in serial[k].items():
                np.testing.assert_array_equal(out[k][key], v, err_msg=f"worker {k} {key} round {rep}")
