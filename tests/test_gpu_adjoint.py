"""The differentiable form: a torch.autograd.Function around the device forward solve and its adjoint,
checked against FP64 central differences (h = 1e-6) of the CPU sparse-LU loss with the same frozen pins,
and against the VJP identity <g, d uv/dp . v> on the forward map itself."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _chart(orc, n=10, seed=11):
    """A small height-field chart: positions, triangles and the reference's border pins."""
    from researchuv_amd import lscm, meshgen

    pos, faces = meshgen.heightfield(n, n, seed=seed)
    pw, fw = lscm.weld(pos, faces, 1e-12)
    seg = lscm.segment(pw, fw, 180.0)
    ch = seg.chart(0)
    s = seg.system(0)
    p3 = pw[ch.vertex_ids].astype(np.float64)
    return p3, ch.tris, s["pin_ids"], s["pin_uv"]


def _autograd_fn():
    import torch

    from researchuv_amd import diff

    class LscmLoss(torch.autograd.Function):
        """J(p) = E(uv(p), p) + |uv(p) - uv*|^2; backward by the adjoint (one PCG with the same SPD matrix)."""

        @staticmethod
        def forward(ctx, p3, target, solver):
            uv, energy = solver.forward(p3.detach().cpu().numpy())
            ctx.save_for_backward(p3, target)
            ctx.solver = solver
            ctx.uv = uv
            loss = energy + float(((uv - target.cpu().numpy()) ** 2).sum())
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            p3, target = ctx.saved_tensors
            g_uv = 2.0 * (ctx.uv - target.cpu().numpy())
            grad = ctx.solver.vjp(p3.detach().cpu().numpy(), ctx.uv, g_uv, energy_weight=1.0)
            return torch.from_numpy(grad) * gout, None, None

    return LscmLoss


@pytest.mark.parametrize("mixed", [False, True])
def test_adjoint_gradient_matches_fp64_central_differences(orc, mixed):
    import torch

    from researchuv_amd import diff

    say(f"adjoint vs central differences mixed={mixed}")
    p3, tris, pin_ids, pin_uv = _chart(orc)
    rng = np.random.default_rng(5)
    target = rng.normal(size=(len(p3), 2)) * 0.1
    solver = diff.ChartSolve(tris, pin_ids, pin_uv, mixed=mixed, tol=1e-13 if not mixed else 1e-12)
    Fn = _autograd_fn()
    pt = torch.tensor(p3, dtype=torch.float64, requires_grad=True)
    loss = Fn.apply(pt, torch.tensor(target), solver)
    loss.backward()
    g = pt.grad.numpy()
    # the CPU-LU loss at the same point agrees with the device forward
    ref = orc.lscm_loss_frozen(p3, tris, pin_ids, pin_uv, target)
    assert ref is not None
    assert abs(ref[0] - loss.item()) <= 1e-9 * abs(ref[0])
    # central differences of the CPU-LU loss on a sample of coordinates (free and pinned vertices)
    h = 1e-6
    idx = rng.choice(p3.size, size=24, replace=False)
    worst = 0.0
    scale = np.abs(g).max()
    for k in idx:
        pp = p3.copy().reshape(-1)
        pm = p3.copy().reshape(-1)
        pp[k] += h
        pm[k] -= h
        fd = (orc.lscm_loss_frozen(pp.reshape(-1, 3), tris, pin_ids, pin_uv, target)[0] -
              orc.lscm_loss_frozen(pm.reshape(-1, 3), tris, pin_ids, pin_uv, target)[0]) / (2 * h)
        worst = max(worst, abs(fd - g.reshape(-1)[k]) / scale)
    say(f"  worst |fd - adjoint| / max|grad| = {worst:.2e}")
    assert worst < 1e-5


def test_vjp_of_the_solve_map_matches_directional_derivative(orc):
    """<g, J v> computed by a central difference of the forward map equals <J^T g, v> from the adjoint."""
    from researchuv_amd import diff

    p3, tris, pin_ids, pin_uv = _chart(orc, n=9, seed=2)
    solver = diff.ChartSolve(tris, pin_ids, pin_uv, tol=1e-13)
    rng = np.random.default_rng(9)
    g = rng.normal(size=(len(p3), 2))
    v = rng.normal(size=p3.shape)
    uv, _ = solver.forward(p3)
    h = 1e-6
    up, _ = solver.forward(p3 + h * v)
    um, _ = solver.forward(p3 - h * v)
    lhs = float((g * (up - um)).sum() / (2 * h))
    rhs = float((solver.vjp(p3, uv, g, energy_weight=0.0) * v).sum())
    say(f"  <g, Jv> = {lhs:.8e}  <J^T g, v> = {rhs:.8e}")
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), 1e-12)


def test_batch_union_matches_per_chart_solves_and_gradients(orc):
    """The differentiable batch (disjoint union of charts, one forward and one adjoint solve): uv and dJ/dp
    equal the per-chart solves, and the gradient matches FP64 central differences of the summed CPU-LU loss."""
    import torch

    from researchuv_amd import diff

    say("differentiable batch of 5 charts")
    charts, p3s = [], []
    for k in range(5):
        p3, tris, pin_ids, pin_uv = _chart(orc, n=8 + k, seed=20 + k)
        charts.append((len(p3), tris, pin_ids, pin_uv))
        p3s.append(p3)
    P = np.concatenate(p3s)
    rng = np.random.default_rng(3)
    target = rng.normal(size=(len(P), 2)) * 0.1
    bs = diff.BatchSolve(charts, tol=1e-13)
    Fn = _autograd_fn()
    pt = torch.tensor(P, dtype=torch.float64, requires_grad=True)
    loss = Fn.apply(pt, torch.tensor(target), bs)
    loss.backward()
    g = pt.grad.numpy()
    uv, _ = bs.forward(P)
    for c, (nv, tris, pin_ids, pin_uv) in enumerate(charts):
        cs = diff.ChartSolve(tris, pin_ids, pin_uv, tol=1e-13)
        pc = bs.split(P)[c]
        uc, _ = cs.forward(pc)
        assert np.abs(uc - bs.split(uv)[c]).max() <= 1e-9 * np.abs(uc).max()
        tc = bs.split(target)[c]
        gc = cs.vjp(pc, uc, 2.0 * (uc - tc), energy_weight=1.0)
        assert np.abs(gc - bs.split(g)[c]).max() <= 1e-7 * np.abs(gc).max()

    def cpu_loss(Pm):
        tot = 0.0
        for c, (nv, tris, pin_ids, pin_uv) in enumerate(charts):
            r = orc.lscm_loss_frozen(bs.split(Pm)[c], tris, pin_ids, pin_uv, bs.split(target)[c])
            tot += r[0]
        return tot

    assert abs(cpu_loss(P) - loss.item()) <= 1e-9 * abs(loss.item())
    h = 1e-6
    worst = 0.0
    for k in rng.choice(P.size, size=16, replace=False):
        pp = P.copy().reshape(-1)
        pm = P.copy().reshape(-1)
        pp[k] += h
        pm[k] -= h
        fd = (cpu_loss(pp.reshape(-1, 3)) - cpu_loss(pm.reshape(-1, 3))) / (2 * h)
        worst = max(worst, abs(fd - g.reshape(-1)[k]) / np.abs(g).max())
    say(f"  worst |fd - adjoint| / max|grad| = {worst:.2e}")
    assert worst < 1e-5


@pytest.mark.parametrize("pc,mixed", [("JACOBI", False), ("AMG", True)])
def test_device_diff_batch_matches_union_solve(orc, pc, mixed):
    """The device-resident differentiable batch (structure built once, batched solves) gives the uv, energy and
    gradient of the union solve, through the same torch.autograd.Function, and a second forward at moved
    positions (same structure) matches a fresh solve there."""
    import torch

    from researchuv_amd import diff
    from researchuv_amd.solver import Precond

    say(f"device differentiable batch precond={pc} mixed={mixed}")
    charts, p3s = [], []
    for k in range(4):
        p3, tris, pin_ids, pin_uv = _chart(orc, n=9 + k, seed=40 + k)
        charts.append((len(p3), tris, pin_ids, pin_uv))
        p3s.append(p3)
    P = np.concatenate(p3s)
    rng = np.random.default_rng(8)
    target = rng.normal(size=(len(P), 2)) * 0.1
    ref = diff.BatchSolve(charts, tol=1e-13)
    db = diff.DiffBatch(P, charts, precond=getattr(Precond, pc), mixed=mixed, tol=1e-12)
    Fn = _autograd_fn()
    for step in range(2):
        Pk = P + step * 1e-3 * rng.normal(size=P.shape)
        pt = torch.tensor(Pk, dtype=torch.float64, requires_grad=True)
        loss = Fn.apply(pt, torch.tensor(target), db)
        loss.backward()
        u_ref, e_ref = ref.forward(Pk)
        g_ref = ref.vjp(Pk, u_ref, 2.0 * (u_ref - target), energy_weight=1.0)
        u_dev, e_dev = db.forward(Pk)
        say(f"  step {step}: uv diff {np.abs(u_dev - u_ref).max():.2e}, grad diff "
            f"{np.abs(pt.grad.numpy() - g_ref).max() / np.abs(g_ref).max():.2e}, iters {db.last_info.iters}")
        assert np.abs(u_dev - u_ref).max() <= 1e-8 * np.abs(u_ref).max()
        assert abs(e_dev - e_ref) <= 1e-9 * abs(e_ref)
        assert np.abs(pt.grad.numpy() - g_ref).max() <= 1e-6 * np.abs(g_ref).max()
    db.close()


def _frozen_system(P, tris, pin_ids, pin_uv):
    """The frozen-pin free block L (cotan weights c_e = w_e / |e|^2, lscm.rs:252-274,610, over the free vertices)
    and its right-hand side (pin terms), assembled with numpy -> (L csr, rhs (nfree, 2), free ids)."""
    import scipy.sparse as sp

    t = np.asarray(tris, np.int64)
    e0 = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
    e1 = np.concatenate([t[:, 1], t[:, 2], t[:, 0]])
    o = np.concatenate([t[:, 2], t[:, 0], t[:, 1]])
    a1, a2 = P[e1] - P[o], P[e0] - P[o]
    cot = np.maximum(np.einsum("ij,ij->i", a1, a2) / np.linalg.norm(np.cross(a1, a2), axis=1), 0.0)
    lo, hi = np.minimum(e0, e1), np.maximum(e0, e1)
    key = lo * len(P) + hi
    uk, inv = np.unique(key, return_inverse=True)
    w = np.bincount(inv, weights=cot)
    i, j = uk // len(P), uk % len(P)
    keep = w > 0.0
    i, j, w = i[keep], j[keep], w[keep]
    c = w / ((P[i] - P[j]) ** 2).sum(axis=1)
    nv = len(P)
    free = np.ones(nv, bool)
    free[np.asarray(pin_ids)] = False
    fid = np.flatnonzero(free)
    pos = -np.ones(nv, np.int64)
    pos[fid] = np.arange(len(fid))
    pin_val = np.zeros((nv, 2))
    pin_val[np.asarray(pin_ids)] = pin_uv
    rows, cols, vals = [], [], []
    rhs = np.zeros((len(fid), 2))
    for a, b in ((i, j), (j, i)):
        fa = free[a]
        rows += [pos[a[fa]]]
        cols += [pos[a[fa]]]
        vals += [c[fa]]
        ff = fa & free[b]
        rows += [pos[a[ff]]]
        cols += [pos[b[ff]]]
        vals += [-c[ff]]
        fp = fa & ~free[b]
        np.add.at(rhs, pos[a[fp]], c[fp, None] * pin_val[b[fp]])
    L = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(len(fid),) * 2)
    return L, rhs, fid


def test_device_diff_batch_128_charts_match_lu_gradient(orc):
    """Four 128 x 128 height-field charts with positions rounded to FP32 (the differentiable workload's charts) in
    the device differentiable batch (device AMG, mixed precision, tol 1e-12): uv against SuperLU, and dJ/dp for
    J = E + |uv - target|^2 against the CPU-LU gradient -- SuperLU for the forward and the adjoint solve (L^T = L)
    and the oracle's gradient through the assembly (orc_lscm_grad_scatter)."""
    import scipy.sparse.linalg as spla

    from researchuv_amd import diff, lscm, meshgen
    from researchuv_amd.solver import Precond

    charts, p3s = [], []
    for k in range(4):
        pos, faces = meshgen.heightfield(128, 128, seed=7000 + k)
        pos = pos.astype(np.float32).astype(np.float64)
        pw, fw = lscm.weld(pos, faces, 1e-12)
        seg = lscm.segment(pw, fw, 180.0)
        ch = seg.chart(0)
        sy = seg.system(0)
        charts.append((len(ch.vertex_ids), ch.tris, sy["pin_ids"], sy["pin_uv"]))
        p3s.append(pw[ch.vertex_ids])
    P = np.concatenate(p3s)
    say(f"4 charts of 128^2 (FP32-rounded positions): {len(P)} vertices")
    target = np.random.default_rng(11).normal(size=(len(P), 2)) * 0.1
    db = diff.DiffBatch(P, charts, precond=Precond.AMG, mixed=True, tol=1e-12)
    uv, e = db.forward(P)
    g = db.vjp(P, uv, 2.0 * (uv - target), energy_weight=1.0)
    db.close()
    uvs, gs, ts = db.split(uv), db.split(g), db.split(target)
    worst_uv, worst_g = 0.0, 0.0
    for c, (nv, tris, pin_ids, pin_uv) in enumerate(charts):
        Pc = p3s[c]
        L, rhs, fid = _frozen_system(Pc, tris, pin_ids, pin_uv)
        lu = spla.splu(L.tocsc())
        uref = np.zeros((nv, 2))
        uref[np.asarray(pin_ids)] = pin_uv
        uref[fid] = lu.solve(rhs)
        lam = np.zeros((nv, 2))
        lam[fid] = lu.solve(np.ascontiguousarray(2.0 * (uref - ts[c])[fid]))
        gref = orc.lscm_grad_scatter(Pc, tris, uref, lam, energy_weight=1.0)
        eu = np.linalg.norm(uvs[c] - uref, axis=1).max() / np.linalg.norm(uref, axis=1).max()
        eg = np.abs(gs[c] - gref).max() / np.abs(gref).max()
        say(f"  chart {c}: uv err {eu:.2e}, grad rel diff {eg:.2e}")
        worst_uv, worst_g = max(worst_uv, eu), max(worst_g, eg)
    assert worst_uv < 1e-9
    assert worst_g < 1e-8


def test_numpy_lu_gradient_equals_oracle_vjp(orc):
    """The SuperLU + scatter gradient used above equals the oracle's all-LU VJP on a small chart."""
    p3, tris, pin_ids, pin_uv = _chart(orc, n=12, seed=31)
    import scipy.sparse.linalg as spla

    rng = np.random.default_rng(1)
    target = rng.normal(size=(len(p3), 2))
    L, rhs, fid = _frozen_system(p3, tris, pin_ids, pin_uv)
    lu = spla.splu(L.tocsc())
    u = np.zeros((len(p3), 2))
    u[np.asarray(pin_ids)] = pin_uv
    u[fid] = lu.solve(rhs)
    lam = np.zeros_like(u)
    lam[fid] = lu.solve(np.ascontiguousarray(2.0 * (u - target)[fid]))
    g1 = orc.lscm_grad_scatter(p3, tris, u, lam, 1.0)
    u2, g2 = orc.lscm_vjp_frozen(p3, tris, pin_ids, pin_uv, 2.0 * (u - target), 1.0)
    assert np.abs(u2 - u).max() <= 1e-12 * np.abs(u).max()
    assert np.abs(g1 - g2).max() <= 1e-10 * np.abs(g2).max()
