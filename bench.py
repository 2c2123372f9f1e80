#!/usr/bin/env python3
"""LSCM solve throughput on one MI355X.

Default workload (`batch`): 4096 open height-field patch charts, vertex counts log-uniform in [64, 65536],
aspect ratio uniform in [1, 8], z = sum of 8 seeded sinusoids (amplitude 0.1, frequencies 1-16), xy jitter
+-0.2 spacing; each patch is welded, segmented at 180 degrees into one chart and pinned on its border by the
reference's border pass, and all charts form one batch.  One step = solving every chart's LSCM system (u and v)
to a true FP64 relative residual ||b - A x|| / ||b|| <= 1e-8 per chart (the mixed solve also until its estimated
forward error is <= 1e-7 max|x|), with the matrix, right-hand sides and solution resident in HBM.

value = LSCM unknowns (2 x free vertices, summed over the charts) solved per second by the mixed-precision
batched solve (FP32 PCG preconditioned by an aggregation AMG built on the device inside every step, within
per-chart FP64 iterative refinement).  Beside it: the same batch in plain
FP64, the plain FP64 drop-in called once per chart as the reference's driver does (on an evenly spaced sample
of the charts), the UV and LSCM-energy error of the mixed solution against a CPU FP64 sparse LU (SuperLU) on a
sample of charts, the step kernel's roofline (timed with HIP events on the solver's stream) and the CPU
oracle's worker-thread sparse-LU path on a bounded sample of the same patch family.

`--workload chart512`: the single 512 x 512 vertex height-field chart; `--workload stretched`: the
ill-conditioned 2048 x 2048 planar grid (column spacing 1 -> 100, 5% flipped diagonals).  Both solve one chart
through the batched engine (a batch of one) with the device-built AMG, the plan drop-in timed beside it.
`--workload diffbatch`: forward + adjoint solves of 256 height-field charts of 128 x 128 (the differentiable
batch).  `--workload metrics`: the post-solve chart metrics (chart_distortion and rectangularize) of the 4096
patch charts on the device, in triangles per second.
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "LSCM unknowns solved/s @ rel. residual 1e-8; SpMV HBM GB/s % peak vs FP64 PCG"
TOL = 1e-8
UV_EST_TOL = 1e-7  # the mixed solves' per-chart forward-error estimate bound (relative to max |x|), library default


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_batch(count=4096, seed=0, threads=16):
    """4096 open grid-patch charts, vertex counts log-uniform in [64, 65536], aspect 1-8, seeded sinusoidal z
    and +-0.2 jitter; each patch its own mesh and chart (segmented at 180 degrees)."""
    from researchuv_amd import batch as B
    from researchuv_amd import meshgen

    meshes = list(meshgen.patch_batch(count, seed=seed))
    lb = B.LscmBatchBuilder(meshes, angle_min_deg=180.0, threads=threads)
    return lb, meshes


def host_threads():
    return max(1, min(16, os.cpu_count() or 1))


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the baseline's record."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_batch(meshes, budget_verts=750_000, vmax=4096):
    """The CPU oracle's worker-thread path (pipeline.rs:171-194 over exec.rs:28-60, each chart one sparse-LU
    factor+solve, sparse.rs:91-265) on the first patches of the batch with at most `vmax` vertices, up to
    `budget_verts` vertices in total (the LU of the largest patches would take minutes each)."""
    import numpy as np

    import oracle

    pos_all, faces_all, nv, picked = [], [], 0, 0
    for k, (pos, faces) in enumerate(meshes):
        if len(pos) > vmax:
            continue
        faces_all.append(faces.astype(np.int64) + sum(len(p) for p in pos_all))
        pos_all.append(pos + np.array([3.0 * k, 0.0, 0.0]))  # disjoint copies: no welding across patches
        nv += len(pos)
        picked += 1
        if nv >= budget_verts:
            break
    P = np.concatenate(pos_all)
    F = np.concatenate(faces_all).astype(np.uint32)
    pw, fw = oracle.weld(P, F, 1e-12)
    ch = oracle.Charts(pw, fw, 180.0)
    unk = 0
    for i in range(len(ch)):
        unk += 2 * ch.system(i)["nfree"]
    w = host_threads()
    t0 = time.perf_counter()
    ch.par_unfold(w, max_iter=1)
    took = time.perf_counter() - t0
    return dict(value=unk / took, unit="unknowns/s", cores=w, kind="port", cpu=cpu_model(),
                sample=f"oracle worker-thread sparse-LU path (one factor+solve per chart) over the first {picked} "
                       f"batch patches with <= {vmax} vertices ({len(ch)} charts, {unk} unknowns), {took:.2f} s "
                       f"on {w} threads; smaller charts than the batch average, so this favours the CPU")


def cpu_baseline_chart(seconds_target=12.0):
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
        ch.unfold(0, max_iter=1)
        took = time.perf_counter() - t0
        unk = 2 * s["nfree"]
        if took > seconds_target / 4 or side >= 160:
            break
        side = int(side * 1.4)
    return dict(value=unk / took, unit="unknowns/s", cores=1, kind="port", cpu=cpu_model(),
                sample=f"oracle sparse-LU unfold (one factor+solve) of one {side}x{side} height-field chart "
                       f"({unk} unknowns), {took:.2f} s on 1 core")


def chart_blocks(a, lb):
    """Per chart: (L as scipy CSR, rhs (n, 2)) from the batch builder's arrays."""
    import numpy as np
    import scipy.sparse as sp

    off = a["chart_row_off"]
    ptr, cols, vals, rhs = a["row_ptr"], a["cols"], a["vals"], a["rhs"]

    def block(c):
        r0, r1 = int(off[c]), int(off[c + 1])
        p = ptr[r0:r1 + 1] - ptr[r0]
        L = sp.csr_matrix((vals[ptr[r0]:ptr[r1]], cols[ptr[r0]:ptr[r1]] - r0, p), shape=(r1 - r0, r1 - r0))
        return L, rhs[2 * r0:2 * r1].reshape(-1, 2)

    return block


def step_profile(bt, args, reps=5):
    """Per-kernel roofline of the AMG-preconditioned step over the whole batch (every chart running): HIP events
    around every launch of `reps` steps (ruv_batch_step_profile), the algorithmic bytes of each launch (nonzeros'
    values and columns, each vector element read or written once, index arrays; stored bytes count the padding
    slots), summed over a step per kernel (all levels of a coarse kernel together).  Outside the timed region."""
    prof = bt.step_profile(reps)
    for e in prof:
        log(f"  step profile: {e['kernel']:12s} level {e['level']:2d} {e['seconds'] * 1e6:8.1f} us "
            f"{e['bytes'] / 1e6:9.1f} MB ({e['stored_bytes'] / 1e6:9.1f} MB stored) -> "
            f"{e['bytes'] / e['seconds'] / 1e9:6.0f} GB/s")
    agg = {}
    for e in prof:
        a = agg.setdefault(e["kernel"], dict(kernel=e["kernel"], launches=0, us=0.0, bytes=0.0, stored_bytes=0.0))
        a["launches"] += 1
        a["us"] += e["seconds"] * 1e6
        a["bytes"] += e["bytes"]
        a["stored_bytes"] += e["stored_bytes"]
    rows = []
    for a in sorted(agg.values(), key=lambda a: -a["us"]):
        gbps = a["bytes"] / (a["us"] * 1e-6) / 1e9
        rows.append(dict(kernel=a["kernel"], launches_per_step=a["launches"], us_per_step=round(a["us"], 1),
                         MB_per_step=round(a["bytes"] / 1e6, 1), GBps=round(gbps), frac=round(gbps * 1e9 / HBM_PEAK, 3),
                         stored_GBps=round(a["stored_bytes"] / (a["us"] * 1e-6) / 1e9)))
    total = sum(r["us_per_step"] for r in rows)
    amg = [r for r in rows if r["kernel"].startswith("kc_")]
    top = amg[0] if amg else None
    log(f"step profile (whole batch): {total:.0f} us per step; top AMG kernel {top}")
    return {"step_profile_whole_batch": {"us_per_step": round(total, 1), "reps": reps, "kernels": rows},
            "amg_top_kernel": None if top is None else
            {"kernel": top["kernel"], "us_per_step": top["us_per_step"], "GBps": top["GBps"], "frac": top["frac"]}}


def run_batch(args):
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch

    from researchuv_amd import solver

    t0 = time.perf_counter()
    lb, meshes = build_batch(args.charts)
    a = lb.arrays(csr=True)
    log(f"batch: {lb.nchart} charts, {lb.nv_total} verts, {2 * lb.nrows} unknowns, nnz(L)={lb.nnz}, "
        f"host build {time.perf_counter() - t0:.1f}s")
    unknowns = 2 * lb.nrows
    off = a["chart_row_off"]
    dev = torch.device("cuda:0")
    b = torch.tensor(a["rhs"], dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()

    if args.calibrate_copy:  # a streaming kernel of known traffic (1 GiB read, 1 GiB written) for counter runs
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    pc = getattr(solver.Precond, args.batch_precond)
    t0 = time.perf_counter()
    bt = lb.device_batch(precond=pc, mixed=True)
    setup_s = time.perf_counter() - t0
    sched = bt.schedule()
    log(f"mixed batch setup {setup_s:.2f}s (sliced ELL + autotune + preconditioner): {sched} {bt.precond_info()}")

    def step():  # one batch: the preconditioner's device setup (AMG hierarchy) and the solve
        bt.setup_precond()
        return bt.solve_dev(b.data_ptr(), x.data_ptr(), 100000, TOL, per_chart=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    bt.kernel_timing(True)  # HIP events around every step-kernel launch of the timed solves
    t0 = time.perf_counter()
    for _ in range(args.steps):
        info, it, res = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ks = bt.kernel_stats()
    bt.kernel_timing(False)
    value = unknowns * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    pinfo = bt.precond_info()
    log(f"mixed batched ({args.batch_precond.lower()}): {ms_per_step:.1f} ms/batch, max inner iters {info.iters}, "
        f"worst true resid {info.true_rel_resid:.2e}, preconditioner setup {pinfo}")
    achieved = ks["bytes"] / ks["seconds"]
    log(f"SpMV over the timed solves: {ks['launches']} launches, {1e6 * ks['seconds'] / ks['launches']:.1f} "
        f"us/launch, {ks['bytes'] / ks['launches'] / 1e9:.3f} GB/launch algorithmic "
        f"({ks['stored_bytes'] / ks['launches'] / 1e9:.3f} GB as stored) -> {achieved / 1e9:.0f} GB/s")
    sec, byt, sto = bt.spmv_time(20, stored=True)
    log(f"SpMV over the whole batch: {sec * 1e3:.3f} ms/launch, {byt / 1e9:.3f} GB algorithmic ({sto / 1e9:.3f} GB "
        f"as stored) -> {byt / sec / 1e9:.0f} GB/s")
    step_prof = step_profile(bt, args) if args.batch_precond == "AMG" and not args.no_step_profile else None
    bt.close()
    if args.no_compare:
        return dict(value=value, ms_per_step=ms_per_step, dtype="f32 inner / f64 refinement",
                    config={"workload": f"batch: {lb.nchart} uneven height-field patch charts", "charts": lb.nchart,
                            "unknowns": unknowns, "tol_true_rel_resid": TOL,
                            "tol_forward_error_estimate": UV_EST_TOL, "schedule": sched},
                    extra={"max_inner_iters": int(info.iters), "precond_setup_in_step": pinfo,
                           "note": "comparison legs skipped (--no-compare)", **(step_prof or {})},
                    roofline=dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                                  frac=achieved / HBM_PEAK, traffic=None), meshes=meshes)

    # the same batch in plain FP64: with the same device AMG (isolates the mixed-precision gain), and with
    # Jacobi (the reference's default preconditioner, lscm.rs:488); both stop on the recursive residual at tol
    bt64a = lb.device_batch(precond=pc, mixed=False)
    x64a = torch.zeros_like(b)

    def step64():
        bt64a.setup_precond()
        return bt64a.solve_dev(b.data_ptr(), x64a.data_ptr(), 100000, TOL, per_chart=True)

    step64()
    torch.cuda.synchronize()
    n64 = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(n64):
        info64a, _, _ = step64()
    torch.cuda.synchronize()
    el64a = (time.perf_counter() - t0) / n64
    bt64a.close()
    log(f"fp64 batched ({args.batch_precond.lower()}): {el64a * 1e3:.1f} ms/batch, max iters {info64a.iters}, worst "
        f"true resid {info64a.true_rel_resid:.2e}")
    bt64 = lb.device_batch(mixed=False)
    x64 = torch.zeros_like(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info64, it64, res64 = bt64.solve_dev(b.data_ptr(), x64.data_ptr(), 100000, TOL, per_chart=True)
    torch.cuda.synchronize()
    el64 = time.perf_counter() - t0
    sec64, byt64 = bt64.spmv_time(10)
    bt64.close()
    log(f"fp64 batched (jacobi): {el64 * 1e3:.1f} ms, max iters {info64.iters}, worst true resid "
        f"{info64.true_rel_resid:.2e}")

    # UV and LSCM-energy error per chart against a CPU FP64 sparse LU (SuperLU) outside the timed region: the
    # 64 largest charts (the ones a residual test bounds least) and 32 evenly spaced ones, for every solve
    sols = {"mixed": x.cpu().numpy().reshape(-1, 2), "fp64_amg": x64a.cpu().numpy().reshape(-1, 2),
            "fp64_jacobi": x64.cpu().numpy().reshape(-1, 2)}
    block = chart_blocks(a, lb)
    sizes = np.diff(off)
    largest = [int(c) for c in np.argsort(-sizes, kind="stable")[:64]]
    spaced = [c for c in range(0, lb.nchart, max(1, lb.nchart // 32)) if c not in largest][:32]
    pick = largest + spaced

    def lu_check(c):
        L, r = block(c)
        xr = spla.splu(L.tocsc()).solve(r)

        def energy(u):  # the quadratic the solve minimises, per coordinate: 1/2 u^T L u - r^T u
            return float(sum(0.5 * u[:, k] @ (L @ u[:, k]) - r[:, k] @ u[:, k] for k in range(2)))

        e_ref = energy(xr)
        sc = max(float(np.linalg.norm(xr, axis=1).max()), 1e-300)
        out = {}
        for name, xs in sols.items():
            xg = xs[off[c]:off[c + 1]]
            out[name] = (float(np.linalg.norm(xg - xr, axis=1).max()) / sc,
                         abs(energy(xg) - e_ref) / max(abs(e_ref), 1e-300))
        return out

    t0 = time.perf_counter()
    with ThreadPoolExecutor(host_threads()) as ex:
        checks = dict(zip(pick, ex.map(lu_check, pick)))
    uvcheck = {}
    for name in sols:
        uv_l = [checks[c][name][0] for c in largest]
        uv_all = [checks[c][name][0] for c in pick]
        uvcheck[name] = {"uv_rel_err_max": max(uv_all), "uv_rel_err_max_largest64": max(uv_l),
                         "uv_rel_err_median_largest64": float(np.median(uv_l)),
                         "charts_above_1e-6": int(sum(e > 1e-6 for e in uv_all)),
                         "energy_rel_err_max": max(checks[c][name][1] for c in pick)}
        log(f"vs SuperLU ({name}): worst uv err {uvcheck[name]['uv_rel_err_max']:.2e} (largest 64: max "
            f"{uvcheck[name]['uv_rel_err_max_largest64']:.2e}, median {uvcheck[name]['uv_rel_err_median_largest64']:.2e}), "
            f"{uvcheck[name]['charts_above_1e-6']} charts above 1e-6, energy rel err "
            f"{uvcheck[name]['energy_rel_err_max']:.2e}")
    log(f"LU checks of {len(pick)} charts ({sizes[largest[-1]]}..{sizes[largest[0]]} rows for the largest 64) "
        f"took {time.perf_counter() - t0:.1f} s")
    # 12 mid-sized charts for the CPU leg's check against the repository's own FP64 sparse LU (the oracle's
    # restatement of sparse.rs:91-265, too slow for the largest charts: 33 s at 16 000 rows)
    mids = [int(c) for c in np.argsort(sizes, kind="stable") if 1000 <= sizes[c] <= 2500]
    oracle_sample = []
    for c in mids[:: max(1, len(mids) // 12)][:12]:
        L, r = block(c)
        r0 = int(off[c])
        oracle_sample.append((L, r, sols["mixed"][r0:r0 + L.shape[0]].copy()))
    xa = sols["mixed"]
    uv_err, en_err = uvcheck["mixed"]["uv_rel_err_max"], uvcheck["mixed"]["energy_rel_err_max"]

    # plain FP64 drop-in, one call per chart as the reference's driver makes (host CSR of A = L (x) I2 in,
    # host solution out; lscm.rs:471-506 -> solver.rs:284-565), on an evenly spaced sample of charts
    gs = solver.GpuSolver.new()
    sample = list(range(0, lb.nchart, max(1, lb.nchart // args.dropin_charts)))
    I2 = sp.identity(2, format="csr")
    prepared = []
    for c in sample:
        L, r = block(c)
        A = sp.kron(L, I2, format="csr")
        A.sort_indices()
        prepared.append((solver.CsrMatrix(A.data, A.indices.astype(np.int32), A.indptr.astype(np.int32)),
                         r.reshape(-1).copy()))
    gs.cg_solve_precond(prepared[0][0], prepared[0][1], 100000, TOL, solver.Precond.JACOBI)  # warm
    t0 = time.perf_counter()
    unk_d = 0
    for A, r in prepared:
        out = gs.cg_solve_precond(A, r, 100000, TOL, solver.Precond.JACOBI)
        assert out is not None, gs.last_error
        unk_d += len(r)
    el_d = time.perf_counter() - t0
    value_d = unk_d / el_d
    log(f"fp64 drop-in per chart: {len(prepared)} charts, {unk_d} unknowns in {el_d:.2f}s -> {value_d:.3e} /s")

    roof = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK,
                traffic=None, kernel="kb_spmv<float, 2, %d, %d, 0>" % (0 if sched["col_bytes"] == 2 else 1, sched["nontemporal"]),
                launches=ks["launches"], bytes_per_launch=ks["bytes"] / ks["launches"],
                stored_bytes_per_launch=ks["stored_bytes"] / ks["launches"],
                achieved_stored=ks["stored_bytes"] / ks["seconds"] / 1e9,
                us_per_launch=1e6 * ks["seconds"] / ks["launches"],
                whole_batch_launch={"us": sec * 1e6, "bytes": byt, "stored_bytes": sto, "GBps": byt / sec / 1e9})
    # counter-derived HBM bytes per launch, from the profile of this same kernel variant (the autotuner picks it)
    for tf in sorted(glob.glob(os.path.join(ROOT, "profiles", f"traffic_batch_{args.batch_precond.lower()}*.json"))):
        tj = json.load(open(tf))
        if tj.get("kernel") == roof["kernel"]:
            roof["traffic"] = tj.get("hbm_bytes_per_launch")
            break
    return dict(
        value=value, ms_per_step=ms_per_step, dtype="f32 inner / f64 refinement",
        config={"workload": f"batch: {lb.nchart} uneven height-field patch charts (64-65536 vertices, log-uniform), "
                            f"{lb.nv_total} vertices, {unknowns} unknowns", "charts": lb.nchart,
                "unknowns": unknowns, "nnz_L": lb.nnz, "precond": args.batch_precond.lower(),
                "tol_true_rel_resid": TOL, "tol_forward_error_estimate": UV_EST_TOL, "schedule": sched},
        extra={"setup_s": setup_s, "precond_setup_in_step": pinfo, "max_inner_iters": int(info.iters),
               "mean_inner_iters": float(np.mean(it)),
               "worst_true_rel_resid": float(info.true_rel_resid),
               "uv_rel_err_vs_cpu_lu": uv_err, "energy_rel_err_vs_cpu_lu": en_err,
               "cpu_lu_check": f"SuperLU per chart on the 64 largest charts ({int(sizes[largest[-1]])}-"
                               f"{int(sizes[largest[0]])} free vertices) and {len(spaced)} evenly spaced ones; "
                               f"uv err = max |uv - uv_lu| / max |uv_lu| per chart (lscm.rs:749-758)",
               "uv_check": uvcheck,
               "fp64_amg_batched": {"value": unknowns / el64a, "ms": 1e3 * el64a, "max_iters": int(info64a.iters),
                                    "worst_true_rel_resid": float(info64a.true_rel_resid)},
               "fp64_batched": {"value": unknowns / el64, "ms": 1e3 * el64, "max_iters": int(info64.iters),
                                "worst_true_rel_resid": float(info64.true_rel_resid),
                                "spmv_GBps": byt64 / sec64 / 1e9, "precond": "jacobi"},
               "fp64_dropin_per_chart": {"value": value_d, "charts": len(prepared), "unknowns": unk_d,
                                         "seconds": el_d},
               "speedup_vs_fp64_amg_batched": value / (unknowns / el64a),
               "speedup_vs_fp64_batched": value / (unknowns / el64),
               "speedup_vs_fp64_dropin_per_chart": value / value_d, **(step_prof or {})},
        roofline=roof, meshes=meshes, oracle_sample=oracle_sample)


def oracle_lu_check(sample):
    """Part of the CPU leg: the mixed batch's solution of each sampled chart, and SciPy's SuperLU (which the bench's
    checks of the largest charts use), against the oracle's FP64 sparse LU (sparse.rs:91-265), the UV measure of
    lscm.rs:749-758."""
    import numpy as np
    import scipy.sparse.linalg as spla

    import oracle

    def one(item):
        L, r, xg = item
        xo = np.stack([oracle.csr_sparse_solve(L.data, L.indices, L.indptr, r[:, k].copy()) for k in range(2)], 1)
        xs = spla.splu(L.tocsc()).solve(r)
        den = np.linalg.norm(xo, axis=1).max()
        return np.linalg.norm(xg - xo, axis=1).max() / den, np.linalg.norm(xs - xo, axis=1).max() / den

    t0 = time.perf_counter()
    with ThreadPoolExecutor(host_threads()) as ex:
        res = list(ex.map(one, sample))
    out = {"charts": len(sample), "rows": [int(min(x[0].shape[0] for x in sample)), int(max(x[0].shape[0] for x in sample))],
           "mixed_uv_rel_err_max": float(max(e[0] for e in res)),
           "superlu_vs_oracle_uv_rel_diff_max": float(max(e[1] for e in res)),
           "seconds": time.perf_counter() - t0}
    log(f"vs the oracle's sparse LU: {out}")
    return out


def build_single(kind, side):
    """One chart: `chart512` = a side x side height field (seed 1); `stretched` = the planar side x side grid
    with column spacing growing geometrically 1 -> 100 and 5% of the quad diagonals flipped (seed 1)."""
    from researchuv_amd import lscm, meshgen

    if kind == "stretched":
        pos, faces = meshgen.stretched_grid(side, stretch=100.0, flip_frac=0.05, seed=1)
    else:
        pos, faces = meshgen.heightfield(side, side, seed=1)
    pw, fw = lscm.weld(pos, faces, 1e-12)
    seg = lscm.segment(pw, fw, 180.0)
    assert len(seg) == 1, f"expected one chart, got {len(seg)}"
    return seg.system(0)


def run_single(args):
    """One large chart through the batched engine (a batch of one) with the device-built AMG, mixed precision;
    beside it the plan drop-in (FP64, the reference's preconditioners) with an iteration cap."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch

    from researchuv_amd import batch as B
    from researchuv_amd import solver

    kind = args.workload
    side = args.side or (512 if kind == "chart512" else 2048)
    t0 = time.perf_counter()
    s = build_single(kind, side)
    nfree = s["nfree"]
    unknowns = 2 * nfree
    log(f"{kind} {side}x{side}: nfree={nfree} unknowns={unknowns} nnz(L)={len(s['l_vals'])} "
        f"host build {time.perf_counter() - t0:.2f}s")
    dev = torch.device("cuda:0")
    b = torch.tensor(s["rhs"], dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    pc = getattr(solver.Precond, args.batch_precond)
    t0 = time.perf_counter()
    bt = B.Batch.from_csr(s["l_ptr"].astype(np.int64), s["l_cols"], s["l_vals"], [0, nfree], precond=pc, mixed=True)
    setup_s = time.perf_counter() - t0

    def step():
        bt.setup_precond()
        return bt.solve_dev(b.data_ptr(), x.data_ptr(), 1_000_000, TOL, per_chart=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    bt.kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        info, it, res = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ks = bt.kernel_stats()
    bt.kernel_timing(False)
    value = unknowns * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    pinfo = bt.precond_info()
    log(f"mixed batch-of-one ({args.batch_precond.lower()}): {ms_per_step:.1f} ms/solve, iters {info.iters}, true "
        f"{info.true_rel_resid:.2e}, {pinfo}")
    achieved = ks["bytes"] / ks["seconds"]
    sched = bt.schedule()
    roof = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK,
                traffic=None, kernel="kb_spmv<float, 2, %d, %d, 0>" % (0 if sched["col_bytes"] == 2 else 1, sched["nontemporal"]),
                launches=ks["launches"], bytes_per_launch=ks["bytes"] / ks["launches"],
                us_per_launch=1e6 * ks["seconds"] / ks["launches"])
    bt.close()
    xa = x.cpu().numpy().reshape(-1, 2)
    extra = {"setup_s": setup_s, "precond_setup_in_step": pinfo, "iters": int(info.iters),
             "true_rel_resid": float(info.true_rel_resid), "schedule": sched}
    # UV / energy error against SuperLU where its factorisation is affordable
    xr = None
    if nfree <= 400_000:
        L = sp.csr_matrix((s["l_vals"], s["l_cols"], s["l_ptr"]), shape=(nfree, nfree))
        r = s["rhs"].reshape(-1, 2)
        lu = spla.splu(L.tocsc())
        xr = np.stack([lu.solve(r[:, 0]), lu.solve(r[:, 1])], axis=1)
        extra["uv_rel_err_vs_cpu_lu"] = float(np.linalg.norm(xa - xr, axis=1).max() / np.linalg.norm(xr, axis=1).max())

        def energy(u):
            return float(sum(0.5 * u[:, k] @ (L @ u[:, k]) - r[:, k] @ u[:, k] for k in range(2)))

        extra["energy_rel_err_vs_cpu_lu"] = abs(energy(xa) - energy(xr)) / abs(energy(xr))
        log(f"vs SuperLU: uv err {extra['uv_rel_err_vs_cpu_lu']:.2e}, energy {extra['energy_rel_err_vs_cpu_lu']:.2e}")
    # the plan drop-in beside it, one call per preconditioner: plain FP64 PCG (the reference's recurrence and
    # recursive-residual test) and, for IC(0), the configuration's FP32 PCG + IC(0) inside FP64 refinement (true
    # residual test); each timed after one warm solve, setup timed apart (plan creation: for IC(0) the structure,
    # levels and colouring on the host, the factor's values on the device) and, for the preconditioners a plan
    # refreshes on the device, the values update alone (refactor_s: the IC(0) factorisation kernels)
    Lp = solver.CsrMatrix(s["l_vals"], s["l_cols"], s["l_ptr"])
    xp = torch.zeros_like(b)
    compare = args.compare if args.compare is not None else (
        ["IC0", "IC0_COLOR", "JACOBI"] if kind == "chart512" else ["JACOBI"])
    for name in compare:
        for mixed in ((False, True) if name in ("IC0", "IC0_COLOR") else (False,)):
            key = f"{'mixed' if mixed else 'fp64'}_plan_{name.lower()}"
            try:
                t0 = time.perf_counter()
                plan = solver.Plan(Lp, 2, getattr(solver.Precond, name), mixed=mixed)
                st = time.perf_counter() - t0
                plan.solve_dev(b.data_ptr(), xp.data_ptr(), args.compare_iters, TOL)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ip = plan.solve_dev(b.data_ptr(), xp.data_ptr(), args.compare_iters, TOL)
                torch.cuda.synchronize()
                el = time.perf_counter() - t0
                rf = None
                if name in ("IC0", "IC0_COLOR", "JACOBI"):
                    t0 = time.perf_counter()
                    plan.update_values(Lp.vals)
                    rf = time.perf_counter() - t0
                plan.close()
                rec = {"value": unknowns / el, "solve_s": el, "setup_s": st, "refactor_s": rf, "iters": int(ip.iters),
                       "refinements": int(ip.refinements), "true_rel_resid": float(ip.true_rel_resid)}
                if xr is not None:
                    xq = xp.cpu().numpy().reshape(-1, 2)
                    rec["uv_rel_err_vs_cpu_lu"] = float(np.linalg.norm(xq - xr, axis=1).max() /
                                                        np.linalg.norm(xr, axis=1).max())
                extra[key] = rec
                log(f"{key}: setup {st:.2f}s (values update {rf}) solve {el * 1e3:.1f} ms iters {ip.iters} true resid "
                    f"{ip.true_rel_resid:.2e} uv err {rec.get('uv_rel_err_vs_cpu_lu', float('nan')):.2e}")
            except Exception as e:  # not converged within the cap, or a failed setup: reported, not fatal
                extra[key] = {"error": str(e)[:200], "iter_cap": args.compare_iters}
                log(f"{key}: {e}")
    for name in ("ic0", "ic0_color"):
        a64, a32 = extra.get(f"fp64_plan_{name}", {}), extra.get(f"mixed_plan_{name}", {})
        if "solve_s" in a64 and "solve_s" in a32:
            extra[f"speedup_mixed_vs_fp64_plan_{name}"] = a64["solve_s"] / a32["solve_s"]
    what = ("one 512x512-vertex height-field chart" if kind == "chart512" else
            f"one {side}x{side}-vertex planar grid, column spacing 1 -> 100 geometric, 5% flipped diagonals")
    return dict(
        value=value, ms_per_step=ms_per_step, dtype="f32 inner / f64 refinement",
        config={"workload": f"{kind}: {what}, {unknowns} unknowns", "unknowns": unknowns,
                "nnz_L": int(len(s["l_vals"])), "precond": args.batch_precond.lower(), "tol_true_rel_resid": TOL,
                "tol_forward_error_estimate": UV_EST_TOL},
        extra=extra, roofline=roof, meshes=None)


def run_diffbatch(args):
    """The differentiable batch: 256 charts of 128 x 128 height-field vertices (positions rounded to FP32),
    loss = the solver's LSCM energy + |uv - uv*|^2 to seeded targets.  One step = device assembly, forward
    batched solve (AMG rebuilt), adjoint batched solve and gather (dJ/dpositions); value counts both solves'
    unknowns.  The gradient is checked against FP64 central differences (h = 1e-6) of the forward map on a few
    coordinates."""
    import numpy as np

    from researchuv_amd import diff, lscm, meshgen
    from researchuv_amd.solver import Precond

    count, side = args.charts if args.charts != 4096 else 256, args.side or 128
    t0 = time.perf_counter()
    charts, p3s = [], []
    for k in range(count):
        pos, faces = meshgen.heightfield(side, side, seed=7000 + k)
        pos = pos.astype(np.float32).astype(np.float64)
        pw, fw = lscm.weld(pos, faces, 1e-12)
        seg = lscm.segment(pw, fw, 180.0)
        ch = seg.chart(0)
        sy = seg.system(0)
        charts.append((len(ch.vertex_ids), ch.tris, sy["pin_ids"], sy["pin_uv"]))
        p3s.append(pw[ch.vertex_ids])
    P = np.concatenate(p3s)
    nfree = sum(int(c[0]) - len(c[2]) for c in charts)
    unknowns = 2 * nfree
    target = np.random.default_rng(11).normal(size=(len(P), 2)) * 0.1
    t1 = time.perf_counter()
    bs = diff.DiffBatch(P, charts, precond=Precond.AMG, mixed=True, tol=1e-10)
    log(f"diffbatch: {count} charts of {side}x{side}, {len(P)} vertices, {unknowns} unknowns, host build "
        f"{t1 - t0:.1f}s, device structure + solver {time.perf_counter() - t1:.1f}s")

    def step():
        uv, e = bs.forward(P)
        g = bs.vjp(P, uv, 2.0 * (uv - target), energy_weight=1.0)
        return uv, e, g

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        uv, e, g = step()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    loss = e + float(((uv - target) ** 2).sum())
    # directional check of the adjoint against central differences of the forward map (loss J)
    rng = np.random.default_rng(5)
    h = 1e-6
    worst = 0.0
    for k in rng.choice(P.size, size=3, replace=False):
        pp = P.copy().reshape(-1)
        pm = P.copy().reshape(-1)
        pp[k] += h
        pm[k] -= h
        up, ep = bs.forward(pp.reshape(-1, 3))
        um, em = bs.forward(pm.reshape(-1, 3))
        jp = ep + float(((up - target) ** 2).sum())
        jm = em + float(((um - target) ** 2).sum())
        worst = max(worst, abs((jp - jm) / (2 * h) - g.reshape(-1)[k]) / np.abs(g).max())
    log(f"diffbatch: {ms_per_step:.1f} ms per forward+adjoint, loss {loss:.6e}, worst fd mismatch {worst:.2e}")
    return dict(
        value=2 * unknowns / (elapsed / args.steps), ms_per_step=ms_per_step, dtype="f32 inner / f64 refinement",
        config={"workload": f"diffbatch: {count} height-field charts of {side}x{side} vertices, forward + adjoint, "
                            f"{unknowns} unknowns per solve", "unknowns": unknowns, "precond": "amg",
                "tol_true_rel_resid": 1e-10},
        extra={"loss": loss, "fd_rel_mismatch_3_coords": worst},
        roofline=None, meshes=None)


def run_metrics(args):
    """The post-solve metrics stage (pipeline.rs:171-182) over the 4096 patch charts: chart_distortion
    (metrics.rs:73-141) and rectangularize (:147-227) on the device, inputs resident in HBM.  The charts come
    from the product's host charting (weld, segment at 180 degrees); the unfold is stood in for by each patch's
    planar (x, y) coordinates, which the metrics treat like any unfold."""
    import ctypes as C

    import numpy as np
    import torch

    from researchuv_amd import _lib, lscm, meshgen
    from researchuv_amd import metrics as M

    t0 = time.perf_counter()
    p3s, tris, uvs, bnds = [], [], [], []
    for pos, faces in meshgen.patch_batch(args.charts, seed=0):
        pw, fw = lscm.weld(pos, faces, 1e-12)
        seg = lscm.segment(pw, fw, 180.0)
        for i in range(len(seg)):
            ch = seg.chart(i)
            vids = np.asarray(ch.vertex_ids, np.int64)
            p3s.append(pw[vids])
            tris.append(np.asarray(ch.tris, np.uint32))
            uvs.append(np.ascontiguousarray(pw[vids][:, :2]))
            bnds.append(np.searchsorted(vids, np.asarray(ch.border, np.int64)))
    n = len(p3s)
    off = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.int64)  # noqa: E731
    v_off, t_off, b_off = off(p3s), off(tris), off(bnds)
    nv, nt, nb = int(v_off[-1]), int(t_off[-1]), int(b_off[-1])
    dev = torch.device("cuda")
    d = {k: torch.from_numpy(np.ascontiguousarray(a)).to(dev) for k, a in [
        ("v_off", v_off), ("t_off", t_off), ("b_off", b_off), ("p3", np.concatenate(p3s)),
        ("uv", np.concatenate(uvs)), ("tris", np.concatenate(tris).view(np.int32)),
        ("bnd", np.concatenate(bnds).astype(np.int64))]}
    met = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    cnt = torch.zeros((n, 2), dtype=torch.int64, device=dev)
    q = torch.zeros((nv, 2), dtype=torch.float64, device=dev)
    ext = torch.zeros((n, 2), dtype=torch.float64, device=dev)
    L = M.lib()
    V = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    log(f"metrics: {n} charts, {nv} vertices, {nt} triangles, {nb} border vertices, host charting "
        f"{time.perf_counter() - t0:.1f}s")

    def step():
        _lib.check(L.ruv_chart_metrics_dev(n, V(d["v_off"]), V(d["p3"]), V(d["uv"]), V(d["t_off"]), V(d["tris"]),
                                           V(met), V(cnt)))
        _lib.check(L.ruv_rectangularize_dev(n, V(d["v_off"]), V(d["uv"]), V(d["b_off"]), V(d["bnd"]), V(q),
                                            V(ext)))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / args.steps
    # algorithmic bytes: every input read once (offsets, positions, uv, triangles, border ids), outputs written
    byt = 3 * 8 * (n + 1) + nv * (24 + 16) + nt * 12 + nb * 8 + n * (24 + 16) + nv * 16 + n * 16
    log(f"metrics: {1e3 * sec:.2f} ms per pass, {nt / sec:.3e} triangles/s, {byt / sec / 1e9:.1f} GB/s "
        f"(parity with the oracle is tests/test_gpu_metrics.py)")
    return dict(
        value=nt / sec, ms_per_step=1e3 * sec, dtype="f64", metric="chart metrics triangles/s (chart_distortion + "
        "rectangularize)", unit="triangles/s",
        config={"workload": f"metrics: {n} patch charts, {nv} vertices, {nt} triangles, {nb} border vertices",
                "charts": n, "triangles": nt},
        extra={},
        roofline=dict(bound="hbm", achieved=byt / sec / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                      frac=byt / sec / HBM_PEAK, traffic=None, kernel="k_chart_distortion + k_rectangularize"),
        meshes=None, cpu_sample=(p3s, tris, uvs, bnds))


def cpu_baseline_metrics(sample, budget=200_000_000):
    """The oracle's chart_distortion + rectangularize, one thread, over the first charts up to `budget`
    triangles."""
    import oracle

    p3s, tris, uvs, bnds = sample
    t0 = time.perf_counter()
    nt = 0
    k = 0
    while k < len(p3s) and nt < budget:
        oracle.chart_distortion(p3s[k], tris[k], uvs[k])
        oracle.rectangularize(uvs[k], bnds[k])
        nt += len(tris[k])
        k += 1
    sec = time.perf_counter() - t0
    return dict(value=nt / sec, unit="triangles/s", cores=1, kind="port", cpu=cpu_model(),
                sample=f"oracle chart_distortion + rectangularize over the first {k} charts ({nt} triangles), "
                       f"{sec:.2f} s on one thread")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch", choices=["batch", "chart512", "stretched", "diffbatch", "metrics"])
    ap.add_argument("--side", type=int, default=0, help="chart512 / stretched: vertices per side (512 / 2048)")
    ap.add_argument("--compare", nargs="*", default=None,
                    help="single chart: plan preconditioners timed beside (chart512: IC0 IC0_COLOR JACOBI, FP64 and "
                         "mixed for IC(0); stretched: JACOBI)")
    ap.add_argument("--compare-iters", type=int, default=20000, help="single chart: iteration cap of the comparisons")
    ap.add_argument("--charts", type=int, default=4096, help="batch workload: number of charts")
    ap.add_argument("--dropin-charts", type=int, default=64, help="batch: charts in the per-chart drop-in sample")
    ap.add_argument("--batch-precond", default="AMG", choices=["JACOBI", "AMG"],
                    help="batch: preconditioner of the mixed batched solve (AMG is built on the device each step)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-compare", action="store_true",
                    help="batch: skip the legs after the timed mixed solves (FP64 batches, LU checks, drop-in loop)")
    ap.add_argument("--no-step-profile", action="store_true",
                    help="batch: skip the whole-batch per-kernel profile of the step (kernel traces of the solves)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="batch: stream 1 GiB through one kernel first (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); every box here has one GPU and this build "
              f"has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    log("building workload", args.workload)
    if args.workload == "batch":
        out = run_batch(args)
    elif args.workload == "diffbatch":
        out = run_diffbatch(args)
    elif args.workload == "metrics":
        out = run_metrics(args)
    else:
        out = run_single(args)
    cpu = None
    if not args.no_cpu:
        if args.workload == "metrics":
            log("cpu baseline (oracle metrics)")
            cpu = cpu_baseline_metrics(out["cpu_sample"])
        else:
            log("cpu baseline (oracle sparse LU)")
            cpu = cpu_baseline_batch(out["meshes"]) if out["meshes"] is not None else cpu_baseline_chart()
            if out.get("oracle_sample"):
                cpu["oracle_lu_check"] = oracle_lu_check(out["oracle_sample"])
        log(f"cpu: {cpu['value']:.0f} {cpu['unit']} ({cpu['sample']})")
    result = {
        "metric": out.get("metric", METRIC),
        "value": out["value"],
        "unit": out.get("unit", "unknowns/s"),
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
