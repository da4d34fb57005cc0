"""Adjoint beamlet-weight gradient (gegs_adjoint) through the C ABI.

The dose is linear in the beamlet weights for fixed random numbers, and the
objective L = sum_v m_v (D_v - P_v)^2 is quadratic in D: the central finite
difference of L with common random numbers is then exact up to rounding, and
the gradient equals A^T dL/dD for the per-beamlet dose-influence matrix A,
which the CPU oracle builds one beamlet at a time."""
import numpy as np
import pytest

from gegs_amd import beamlets, native, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def _objective(ct, d):
    ph = ct.phantom
    xc = 0.5 * (ph.xb[:-1] + ph.xb[1:])
    yc = 0.5 * (ph.yb[:-1] + ph.yb[1:])
    zc = 0.5 * (ph.zb[:-1] + ph.zb[1:])
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).reshape(-1)
    target = r < 3.0
    oar = (r >= 4.0) & (r < 5.0)
    m = (target | oar).astype(np.float64)
    p = np.where(target, 1.2 * d[target].mean(), 0.0)
    return m, p


def _loss(d, m, p):
    return float(np.sum(m * (d - p) ** 2))


@pytest.fixture(scope="module")
def ct():
    return workloads.ct()


def test_adjoint_matches_central_differences(gpu_ctx, ct):
    rng = np.random.default_rng(4)
    n_per = 40000
    sid = ct.field.source_to_isocenter
    fields = [workloads.make_field(True, sid, g) for g in (0.0, 72.0)]
    src = ct.source(n_per * len(fields), 44)
    gpu_ctx.load_workload(ct)
    gpu_ctx.set_source(src)
    plan = beamlets.BeamletPlan(gpu_ctx, fields, n_per, n_side=20, pitch=0.5, seed=9, n_batches=2, src=src)
    w = rng.uniform(0.5, 1.5, plan.n_beamlets)
    d = plan.forward(w)
    m, p = _objective(ct, d)
    g = plan.adjoint(2.0 * m * (d - p))
    hit = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
    picks = rng.choice(hit, 20, replace=False)
    h = 0.1
    for b in picks:
        wp, wm = w.copy(), w.copy()
        wp[b] += h
        wm[b] -= h
        fd = (_loss(plan.forward(wp), m, p) - _loss(plan.forward(wm), m, p)) / (2 * h)
        assert abs(fd - g[b]) <= 1e-2 * abs(g[b]), (b, fd, g[b])
        assert abs(fd - g[b]) <= 1e-4 * abs(g[b]) + 1e-9 * np.abs(g).max(), (b, fd, g[b])
    gpu_ctx.set_beamlets(None)


def test_adjoint_matches_influence_matrix(gpu_ctx, ct):
    n = 4000
    src = ct.source(n, 45)
    b = beamlets.beamlet_index(src, ct.field, 4, 2.5)
    order = np.argsort(np.where(b >= 0, b, 16), kind="stable")
    src = src[order]
    gpu_ctx.load_workload(ct)
    gpu_ctx.set_source(src)
    plan = beamlets.BeamletPlan(gpu_ctx, [ct.field], n, n_side=4, pitch=2.5, seed=13, n_batches=2, src=src)
    bl = plan.beamlet
    # influence matrix on the CPU: the oracle transports each beamlet's
    # histories (contiguous after sorting) at unit weight
    A = np.zeros((ct.phantom.nvox, plan.n_beamlets + 1))
    for k in range(plan.n_beamlets + 1):
        idx = np.flatnonzero(bl == k)
        if len(idx) == 0:
            continue
        lo, hi = idx[0], idx[-1] + 1
        assert hi - lo == len(idx)
        A[:, k] = oracle.run(ct, src[lo:hi], seed=13, hist_offset=lo)["dose"]
    w = np.random.default_rng(5).uniform(0.5, 1.5, plan.n_beamlets)
    d_ref = A @ np.append(w, 1.0)
    m, p = _objective(ct, d_ref)
    dLdD = 2.0 * m * (d_ref - p)
    g_ref = A[:, :plan.n_beamlets].T @ dLdD
    d = plan.forward(w)
    assert np.allclose(d, d_ref, rtol=1e-5, atol=1e-9 * d_ref.max())
    g = plan.adjoint(dLdD)
    assert np.allclose(g, g_ref, rtol=1e-9, atol=1e-12 * np.abs(g_ref).max())
    gpu_ctx.set_beamlets(None)


def test_torch_autograd_function(gpu_ctx, ct):
    import torch
    n_per = 20000
    src = ct.source(n_per, 46)
    gpu_ctx.load_workload(ct)
    gpu_ctx.set_source(src)
    plan = beamlets.BeamletPlan(gpu_ctx, [ct.field], n_per, n_side=20, pitch=0.5, seed=3, n_batches=2, src=src)
    w0 = np.random.default_rng(6).uniform(0.5, 1.5, plan.n_beamlets)
    w = torch.tensor(w0, dtype=torch.float32, device="cuda", requires_grad=True)
    d = beamlets.BeamletDose.apply(w, plan)
    m_np, p_np = _objective(ct, d.detach().cpu().numpy())
    m = torch.from_numpy(m_np).cuda()
    p = torch.from_numpy(p_np).cuda()
    loss = (m * (d - p) ** 2).sum()
    loss.backward()
    g_direct = plan.adjoint(2.0 * m_np * (d.detach().cpu().numpy() - p_np))
    assert w.grad is not None and w.grad.shape == w.shape
    assert np.allclose(w.grad.cpu().numpy(), g_direct.astype(np.float32), rtol=1e-5, atol=1e-6 * np.abs(g_direct).max())
    # the device-resident forward gives the host path's dose
    d_host = plan.forward(w0)
    assert d.device.type == "cuda" and d.dtype == torch.float64
    assert np.allclose(d.detach().cpu().numpy(), d_host, rtol=1e-12, atol=1e-12 * d_host.max())
    gpu_ctx.set_beamlets(None)


def test_torch_autograd_plan_queue(gpu_ctx, ct):
    """Three plans of one or two fields in one BeamletDose: dose (3, nvox) on
    the device, and the gradient of a sum of per-plan objectives equals the
    host-side adjoint with one dL/dD row per plan."""
    import torch
    sid = ct.field.source_to_isocenter
    f = [workloads.make_field(True, sid, g) for g in (0.0, 90.0, 180.0, 270.0)]
    n = 15000
    jobs = [dict(field=f[0], plan=0, first=0, count=n), dict(field=f[1], plan=1, first=n, count=n),
            dict(field=f[2], plan=1, first=2 * n, count=n), dict(field=f[3], plan=2, first=3 * n, count=n)]
    src = ct.source(4 * n, 48)
    gpu_ctx.load_workload(ct)
    gpu_ctx.set_source(src)
    q = beamlets.BeamletPlan.from_jobs(gpu_ctx, jobs, 3, src, n_side=20, pitch=0.5, seed=5, n_batches=2)
    assert q.n_beamlets == 1600 and list(q.plan_beamlets(1)) == list(range(400, 1200))
    w0 = np.random.default_rng(7).uniform(0.5, 1.5, q.n_beamlets)
    w = torch.tensor(w0, dtype=torch.float32, device="cuda", requires_grad=True)
    d = beamlets.BeamletDose.apply(w, q)
    assert tuple(d.shape) == (3, gpu_ctx.nvox)
    dn = d.detach().cpu().numpy()
    ms, ps = zip(*(_objective(ct, dn[k]) for k in range(3)))
    m = torch.from_numpy(np.stack(ms)).cuda()
    p = torch.from_numpy(np.stack(ps)).cuda()
    ((m * (d - p) ** 2).sum()).backward()
    g_host = q.adjoint(2.0 * np.stack(ms) * (dn - np.stack(ps)))
    assert np.allclose(w.grad.cpu().numpy(), g_host.astype(np.float32), rtol=1e-5, atol=1e-6 * np.abs(g_host).max())
    gpu_ctx.set_beamlets(None)


def test_phase_space_weight_gradient(gpu_ctx, ct):
    """dL/dwt of single source records (each history its own beamlet) against
    central differences of L when one record's weight changes, with common
    random numbers (the dose is linear in each record's weight)."""
    n = 20000
    src = ct.source(n, 47)
    gpu_ctx.load_workload(ct)
    gpu_ctx.set_source(src)
    jobs = [dict(field=ct.field, plan=0, first=0, count=n)]
    gpu_ctx.set_beamlets(None)
    d = gpu_ctx.run_batched(jobs, 1, seed=8, n_batches=2)["dose"][0]
    m, p = _objective(ct, d)
    g = beamlets.phase_space_weight_gradient(gpu_ctx, jobs, 1, src, 8, 2.0 * m * (d - p), n_batches=2)
    assert g.shape == (n,)
    gpu_ctx.set_beamlets(None)
    rng = np.random.default_rng(9)
    picks = rng.choice(np.flatnonzero(np.abs(g) > 1e-2 * np.abs(g).max()), 6, replace=False)
    for i in picks:
        h = 0.25 * float(src["wt"][i])
        ls = []
        for sgn in (1.0, -1.0):
            s2 = src.copy()
            s2["wt"][i] = np.float32(float(src["wt"][i]) + sgn * h)
            gpu_ctx.set_source(s2)
            ls.append(_loss(gpu_ctx.run_batched(jobs, 1, seed=8, n_batches=2)["dose"][0], m, p))
        fd = (ls[0] - ls[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-4 * abs(g[i]) + 1e-9 * np.abs(g).max(), (i, fd, g[i])
    gpu_ctx.set_source(src)
