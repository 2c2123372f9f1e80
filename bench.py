#!/usr/bin/env python3
"""LSCM solve throughput on one MI355X.

Workload (default, `chart512`): the single-large-chart configuration -- a 512 x 512 vertex height field over
the unit square (z = sum of 8 sinusoids, amplitude 0.1, seeded frequencies 1-16, xy jitter +-0.2 spacing),
welded and segmented at 180 degrees into one chart, pinned on its border by the reference's border pass.
One step = one full solve of the chart's LSCM system (both u and v) to a true FP64 relative residual
<= 1e-8, with L, the right-hand side and the solution resident in HBM.

value = LSCM unknowns (2 * free vertices) solved per second by the mixed-precision solve (FP32 Jacobi-PCG
inside FP64 iterative refinement).  The plain FP64 Jacobi-PCG on the same system is timed beside it.
roofline = the FP32 SpMV kernel the PCG iterations launch, timed live with HIP events on the solver's
stream; cpu_baseline = the CPU oracle's sparse-LU unfold (the reference's CPU path) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_chart512(seed=1, n=512):
    from researchuv_amd import lscm, meshgen

    pos, faces = meshgen.heightfield(n, n, seed=seed)
    pw, fw = lscm.weld(pos, faces, 1e-12)
    seg = lscm.segment(pw, fw, 180.0)
    assert len(seg) == 1, f"expected one chart, got {len(seg)}"
    return seg.system(0)


def build_batch(count=4096, seed=0, threads=16):
    """Config: 4096 open grid-patch charts, vertex counts log-uniform in [64, 65536], aspect 1-8, seeded
    sinusoidal z and +-0.2 jitter; each patch its own mesh and chart (segmented at 180 degrees)."""
    from researchuv_amd import batch as B
    from researchuv_amd import meshgen

    meshes = list(meshgen.patch_batch(count, seed=seed))
    lb = B.LscmBatchBuilder(meshes, angle_min_deg=180.0, threads=threads)
    del meshes
    return lb


def cpu_baseline(seconds_target=12.0):
    """The CPU oracle's sparse-LU unfold (lscm.rs:512-727 with SolverBackend::Cpu, sparse.rs:91-265) on a
    height-field chart of the same family, sized to take about `seconds_target` on one core."""
    import oracle
    from researchuv_amd import meshgen

    side, took, unk = 24, 0.0, 0
    while True:
        pos, faces = meshgen.heightfield(side, side, seed=1)
        pw, fw = oracle.weld(pos, faces, 1e-12)
        ch = oracle.Charts(pw, fw, 180.0)
        s = ch.system(0)
        t0 = time.perf_counter()
        ch.unfold(0)
        took = time.perf_counter() - t0
        unk = 2 * s["nfree"]
        if took > seconds_target / 4 or side >= 160:
            break
        side = int(side * 1.4)
    return dict(value=unk / took, unit="unknowns/s", cores=1, kind="port",
                sample=f"oracle sparse-LU unfold of one {side}x{side} height-field chart ({unk} unknowns), "
                       f"2 factor+solve passes as the reference driver does, {took:.2f} s on 1 core")


def run_chart512(args):
    import numpy as np
    import torch

    from researchuv_amd import solver

    t0 = time.perf_counter()
    s = build_chart512()
    setup_host = time.perf_counter() - t0
    nfree = s["nfree"]
    unknowns = 2 * nfree
    L = solver.CsrMatrix(s["l_vals"], s["l_cols"], s["l_ptr"])
    log(f"nfree={nfree} unknowns={unknowns} nnz(L)={len(s['l_vals'])} host setup {setup_host:.2f}s")

    dev = torch.device("cuda:0")
    b = torch.tensor(s["rhs"], dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    x64 = torch.zeros_like(b)
    torch.cuda.synchronize()
    tol = 1e-8
    max_iter = 1_000_000

    plan = solver.Plan(L, 2, solver.Precond.JACOBI, mixed=True)
    for _ in range(args.warmup):
        plan.solve_dev(b.data_ptr(), x.data_ptr(), max_iter, tol)
    torch.cuda.synchronize()
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos.append(plan.solve_dev(b.data_ptr(), x.data_ptr(), max_iter, tol))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = unknowns * args.steps / elapsed
    info = infos[-1]
    log(f"mixed: {ms_per_step:.2f} ms/solve, inner iters {info.iters}, refinements {info.refinements}, "
        f"true rel residual {info.true_rel_resid:.2e}")

    # plain FP64 PCG (the drop-in) on the same system, recursive tol 1e-8
    plan64 = solver.Plan(L, 2, solver.Precond.JACOBI, mixed=False)
    plan64.solve_dev(b.data_ptr(), x64.data_ptr(), max_iter, tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        info64 = plan64.solve_dev(b.data_ptr(), x64.data_ptr(), max_iter, tol)
    torch.cuda.synchronize()
    el64 = (time.perf_counter() - t0) / args.fp64_steps
    value64 = unknowns / el64
    log(f"fp64: {1e3 * el64:.2f} ms/solve, iters {info64.iters}, true rel residual {info64.true_rel_resid:.2e}")

    # agreement of the two solutions in UV terms (lscm.rs:749-758 metric)
    xa = x.cpu().numpy().reshape(-1, 2)
    xb = x64.cpu().numpy().reshape(-1, 2)
    uv_diff = float(np.linalg.norm(xa - xb, axis=1).max() / np.linalg.norm(xb, axis=1).max())

    # roofline of the dominant kernel (FP32 SpMV over L, two right-hand sides)
    import ctypes as C

    from researchuv_amd import _lib

    sec = C.c_double(0)
    byt = C.c_double(0)
    _lib.check(_lib.lib().ruv_plan_spmv_time(plan.h, 200, C.byref(sec), C.byref(byt)))
    achieved = byt.value / sec.value
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_chart512.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("spmv_hbm_bytes_per_launch")
    roof = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK,
                traffic=traffic, kernel="k_spmv_pq<float,2>", bytes_per_launch=byt.value,
                us_per_launch=sec.value * 1e6)
    log(f"spmv: {sec.value * 1e6:.2f} us/launch, {byt.value / 1e6:.2f} MB algorithmic -> {achieved / 1e9:.0f} GB/s")

    return dict(
        value=value, ms_per_step=ms_per_step, dtype="f32 inner / f64 refinement",
        config={"workload": "chart512: one 512x512-vertex height-field chart, 520200 unknowns",
                "unknowns": unknowns, "nnz_L": int(len(s["l_vals"])), "precond": "jacobi",
                "tol_true_rel_resid": tol},
        extra={"inner_iters": int(info.iters), "refinements": int(info.refinements),
               "true_rel_resid": float(info.true_rel_resid),
               "fp64_pcg": {"value": value64, "ms_per_solve": 1e3 * el64, "iters": int(info64.iters),
                            "true_rel_resid": float(info64.true_rel_resid)},
               "speedup_vs_fp64": value / value64, "uv_rel_diff_mixed_vs_fp64": uv_diff},
        roofline=roof)


def run_batch(args):
    import numpy as np
    import torch

    t0 = time.perf_counter()
    lb = build_batch(args.charts)
    a = lb.arrays()
    log(f"batch: {lb.nchart} charts, {lb.nv_total} verts, {2 * lb.nrows} unknowns, nnz(L)={lb.nnz}, "
        f"host build {time.perf_counter() - t0:.1f}s")
    unknowns = 2 * lb.nrows
    tol = 1e-8
    dev = torch.device("cuda:0")
    b = torch.tensor(a["rhs"], dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    bt = lb.device_batch(mixed=True)
    for _ in range(args.warmup):
        bt.solve_dev(b.data_ptr(), x.data_ptr(), 16384, tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        info, it, res = bt.solve_dev(b.data_ptr(), x.data_ptr(), 16384, tol, per_chart=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = unknowns * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"mixed batched: {ms_per_step:.1f} ms/batch, max inner iters {info.iters}, worst true resid "
        f"{info.true_rel_resid:.2e}")
    sec, byt = bt.spmv_time(20)
    achieved = byt / sec
    log(f"batched spmv: {sec * 1e3:.3f} ms/launch, {byt / 1e9:.2f} GB -> {achieved / 1e9:.0f} GB/s")
    bt.close()
    # FP64 comparison 1: the same batch in plain FP64 (recursive-residual test at tol)
    bt64 = lb.device_batch(mixed=False)
    x64 = torch.zeros_like(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info64, it64, res64 = bt64.solve_dev(b.data_ptr(), x64.data_ptr(), 16384, tol, per_chart=True)
    torch.cuda.synchronize()
    el64 = time.perf_counter() - t0
    bt64.close()
    xa = x.cpu().numpy().reshape(-1, 2)
    xb = x64.cpu().numpy().reshape(-1, 2)
    off = a["chart_row_off"]
    worst = 0.0
    for c in range(lb.nchart):
        da = xa[off[c]:off[c + 1]]
        db = xb[off[c]:off[c + 1]]
        if len(db):
            worst = max(worst, float(np.linalg.norm(da - db, axis=1).max() / max(np.linalg.norm(db, axis=1).max(), 1e-300)))
    log(f"fp64 batched: {el64 * 1e3:.1f} ms, worst true resid {info64.true_rel_resid:.2e}; mixed vs fp64 uv diff {worst:.2e}")
    roof = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK,
                traffic=None, kernel="kb_spmv<float,2>", bytes_per_launch=byt, us_per_launch=sec * 1e6)
    tf = os.path.join(ROOT, "profiles", "traffic_batch.json")
    if os.path.exists(tf):
        roof["traffic"] = json.load(open(tf)).get("spmv_hbm_bytes_per_launch")
    return dict(
        value=value, ms_per_step=ms_per_step, dtype="f32 inner / f64 refinement",
        config={"workload": f"batch: {lb.nchart} uneven height-field patch charts, {lb.nv_total} vertices, "
                            f"{unknowns} unknowns", "charts": lb.nchart, "unknowns": unknowns, "nnz_L": lb.nnz,
                "precond": "jacobi", "tol_true_rel_resid": tol},
        extra={"max_inner_iters": int(info.iters), "mean_inner_iters": float(np.mean(it)),
               "worst_true_rel_resid": float(info.true_rel_resid),
               "fp64_batched": {"value": unknowns / el64, "ms": 1e3 * el64, "max_iters": int(info64.iters),
                                "worst_true_rel_resid": float(info64.true_rel_resid)},
               "speedup_vs_fp64_batched": (unknowns / elapsed * args.steps) / (unknowns / el64),
               "uv_rel_diff_mixed_vs_fp64": worst},
        roofline=roof)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="chart512", choices=["chart512", "batch"])
    ap.add_argument("--charts", type=int, default=4096, help="batch workload: number of charts")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--fp64-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); every box here has one GPU and this build "
              f"has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from researchuv_amd import solver

    log("building workload", args.workload)
    if args.workload == "batch":
        out = run_batch(args)
    else:
        out = run_chart512(args)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle sparse LU)")
        cpu = cpu_baseline()
        log(f"cpu: {cpu['value']:.0f} unknowns/s ({cpu['sample']})")

    result = {
        "metric": "LSCM unknowns solved/s @ rel. residual 1e-8 (mixed FP32 PCG + FP64 refinement)",
        "value": out["value"],
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": out["dtype"],
        "data": "synthetic (seeded height fields)",
        "config": out["config"],
    }
    result.update(out["extra"])
    result["roofline"] = out["roofline"]
    result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
