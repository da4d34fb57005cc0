"""Full-size workloads on the GPU.

The 60^3 water phantom with 1e6 histories and the 128x128x100 CT with 1e7
histories are compared with the oracle history by history (the oracle runs
on up to 16 host threads).  Where the oracle would take too long, properties
that hold at any size are checked instead: the event-queued loop (one field,
or all fields of a plan in one loop) and the plain per-history kernel
transport the same histories with the same Philox streams, so every integer
counter (photon and electron steps, region gathers, deposits, pops per event
type) is identical, the doses agree up to the order of the FP64 additions,
and no particle is lost to a full bank: the 128x128x100 CT with 1e7 histories
in 10 batches, and the 5-field MLC plan on the same CT with 2e7 phase-space
records (plain: five per-field runs; batched: one run of the five fields)."""
import copy
import os

import numpy as np
import pytest

from gegs_amd import native, workloads
from oracle import oracle

THREADS = max(1, min(16, os.cpu_count() or 1))

pytestmark = pytest.mark.gpu

INT_COUNTERS = ("photon_steps", "electron_steps", "region_gathers", "deposits", "histories")


def _same_counters(a, b):
    for k in INT_COUNTERS:
        assert a[k] == b[k], (k, a[k], b[k])
    assert list(a["interactions"]) == list(b["interactions"])
    assert a["overflow"] == 0 and b["overflow"] == 0


def _sum_counters(acc, c):
    if acc is None:
        return {k: (list(v) if isinstance(v, (list, tuple)) else v) for k, v in c.items()}
    for k, v in c.items():
        if isinstance(v, (list, tuple)):
            acc[k] = [x + y for x, y in zip(acc[k], v)]
        else:
            acc[k] += v
    return acc


def _close(d1, d2):
    scale = float(np.max(np.abs(d1)))
    assert scale > 0
    np.testing.assert_allclose(d1, d2, rtol=1e-10, atol=1e-12 * scale)


def _per_history_vs_oracle(wl, n, src_seed, seed, method):
    src = wl.source(n, src_seed)
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        g = ctx.run(seed=seed, n_batches=10, method=method, per_history=True)
    finally:
        ctx.close()
    o = oracle.run(wl, src, seed=seed, per_history=True, threads=THREADS)
    assert np.array_equal(g["steps"], o["steps"])
    if method == "plain":
        assert np.array_equal(g["edep"], o["edep"])
    else:
        # a history's particles add their sums to its tally in queue order
        np.testing.assert_allclose(g["edep"], o["edep"], rtol=1e-12, atol=1e-12)
    c = o["counters"]
    assert g["counters"]["photon_steps"] == c.photon_steps
    assert g["counters"]["electron_steps"] == c.electron_steps
    assert g["counters"]["interactions"] == list(c.interactions)
    _close(o["dose"], g["dose"])


def test_water_60cubed_1e6_plain_matches_oracle_per_history():
    wl = workloads.water()
    assert (wl.phantom.nx, wl.phantom.ny, wl.phantom.nz) == (60, 60, 60)
    _per_history_vs_oracle(wl, 1_000_000, 1, 1234, "plain")


def test_ct_1e7_event_matches_oracle_per_history():
    _per_history_vs_oracle(workloads.ct(), 10_000_000, 7, 4321, "event")


def test_ct_1e7_event_equals_plain():
    wl = workloads.ct()
    assert (wl.phantom.nx, wl.phantom.ny, wl.phantom.nz) == (128, 128, 100)
    n = 10_000_000
    src = wl.source(n, 2024)
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        p = ctx.run(seed=99, n_batches=10, method="plain")
        e = ctx.run(seed=99, n_batches=10, method="event")
    finally:
        ctx.close()
    assert p["counters"]["histories"] == n
    _same_counters(p["counters"], e["counters"])
    # about 170 photon steps per history on this phantom
    assert 150 * n < p["counters"]["photon_steps"] < 190 * n
    _close(p["dose"], e["dose"])
    _close(p["dose2"], e["dose2"])


def test_mlc_plan_2e7_batched_equals_per_field_plain():
    plan = workloads.mlc_plan()
    n = 20_000_000
    src = plan.source(n, 3)
    jobs = plan.jobs(n)
    assert len(jobs) == 5 and sum(j["count"] for j in jobs) == n
    ctx = native.Context(0)
    try:
        ctx.load_workload(plan.workload)
        ctx.set_source(src)
        ctx.set_leaf_positions(plan.leaf_positions)
        b = ctx.run_batched(jobs, 1, seed=41, n_batches=10)
        dose, cnt = None, None
        for jb in jobs:
            ctx.set_leaf_positions(plan.leaf_positions[jb["leaf_set"]])
            ctx.set_field(jb["field"])
            r = ctx.run(seed=41, n_batches=10, method="plain", first=jb["first"], count=jb["count"])
            dose = r["dose"] if dose is None else dose + r["dose"]
            cnt = _sum_counters(cnt, r["counters"])
    finally:
        ctx.close()
    _same_counters(cnt, b["counters"])
    _close(dose, b["dose"][0])


def test_adjoint_128cube_1e7_matches_central_differences():
    """5 fields x 400 beamlets (20 x 20 of 5 mm), 1e7 histories on the 128^3
    CT, w ~ U(0.5, 1.5): the replayed gradient of L = sum m (D - P)^2 against
    central differences with common random numbers on 20 beamlets (exact up
    to rounding: D is linear in w for fixed random numbers, L quadratic)."""
    from gegs_amd import beamlets
    wl = workloads.ct_cube()
    assert (wl.phantom.nx, wl.phantom.ny, wl.phantom.nz) == (128, 128, 128)
    sid = wl.field.source_to_isocenter
    fields = [workloads.make_field(True, sid, g) for g in workloads.C3_GANTRY]
    n_per = 2_000_000
    src = wl.source(n_per * len(fields), 4)
    rng = np.random.default_rng(4)
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        plan = beamlets.BeamletPlan(ctx, fields, n_per, n_side=20, pitch=0.5, seed=4, n_batches=2, src=src)
        assert plan.n_beamlets == 2000
        w = rng.uniform(0.5, 1.5, plan.n_beamlets)
        d = plan.forward(w)
        m, p = workloads.objective(wl.phantom, d)
        g = plan.adjoint(2.0 * m * (d - p))
        hit = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
        h = 0.1
        for b in rng.choice(hit, 20, replace=False):
            wp, wm = w.copy(), w.copy()
            wp[b] += h
            wm[b] -= h
            lp = float(np.sum(m * (plan.forward(wp) - p) ** 2))
            lm = float(np.sum(m * (plan.forward(wm) - p) ** 2))
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[b]) <= 1e-4 * abs(g[b]) + 1e-9 * np.abs(g).max(), (b, fd, g[b])
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def plan_queue():
    return workloads.multi_plan()


def test_plan_queue_geometry_matches_oracle(plan_queue):
    """The 256 x 256 x 180 CT at 2 mm: the first field of every one of the 16
    plans, 300 histories each, in one batched run, history by history
    against the oracle."""
    mp = plan_queue
    wl = mp.workload
    src = wl.source(300 * len(mp.plans), 55)
    jobs = [dict(field=fs[0], plan=k, first=300 * k, count=300) for k, fs in enumerate(mp.plans)]
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        b = ctx.run_batched(jobs, len(mp.plans), seed=12, n_batches=2, per_history=True)
    finally:
        ctx.close()
    assert b["counters"]["overflow"] == 0
    for jb in jobs:
        one = copy.copy(wl)
        one.field = jb["field"]
        sl = slice(jb["first"], jb["first"] + jb["count"])
        o = oracle.run(one, src[sl], seed=12, hist_offset=jb["first"], per_history=True, threads=THREADS)
        assert np.array_equal(b["steps"][sl], o["steps"])
        np.testing.assert_allclose(b["edep"][sl], o["edep"], rtol=1e-12, atol=1e-12)
        _close(o["dose"], b["dose"][jb["plan"]])


def test_plan_queue_1e8_batched_equals_per_field_plain(plan_queue):
    """16 plans of 1-9 fields, budgets 1e6-2e7 summing to 1e8 histories, on
    the 256 x 256 x 180 CT: one batched run of all 93 fields against the plain
    kernel field after field."""
    mp = plan_queue
    src = mp.source(5)
    jobs = mp.jobs()
    assert sum(j["count"] for j in jobs) == 100_000_000
    ctx = native.Context(0)
    try:
        ctx.load_workload(mp.workload)
        ctx.set_source(src)
        del src
        b = ctx.run_batched(jobs, len(mp.plans), seed=77, n_batches=10)
        cnt = None
        dose = np.zeros_like(b["dose"])
        for jb in jobs:
            ctx.set_field(jb["field"])
            r = ctx.run(seed=77, n_batches=10, method="plain", first=jb["first"], count=jb["count"])
            dose[jb["plan"]] += r["dose"]
            cnt = _sum_counters(cnt, r["counters"])
    finally:
        ctx.close()
    _same_counters(cnt, b["counters"])
    for k in range(len(mp.plans)):
        _close(dose[k], b["dose"][k])


def test_plan_queue_per_plan_adjoint(plan_queue):
    """The 16-plan queue on the 256 x 256 x 180 CT, 20 x 20 beamlets of 5 mm
    per field (every field of every plan), history budgets scaled to 1/100:
    one replay of all plans with one dL/dD row per plan (target sphere and
    OAR shell of each plan's own dose).  Each plan's gradient equals the
    adjoint of that plan's fields alone, and central differences of its
    objective with common random numbers on two beamlets agree."""
    from gegs_amd import beamlets
    mp = plan_queue
    wl = mp.workload
    jobs, first = [], 0
    for jb in mp.jobs():
        n = max(2000, jb["count"] // 100)
        jobs.append(dict(field=jb["field"], plan=jb["plan"], first=first, count=n))
        first += n
    src = wl.source(first, 61)
    n_plans = len(mp.plans)
    rng = np.random.default_rng(12)
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        print("plan-queue adjoint: %d jobs, %d plans, %d histories" % (len(jobs), n_plans, first), flush=True)
        q = beamlets.BeamletPlan.from_jobs(ctx, jobs, n_plans, src, n_side=20, pitch=0.5, seed=31, n_batches=2)
        assert q.n_beamlets == 400 * len(jobs)
        w = rng.uniform(0.5, 1.5, q.n_beamlets)
        d = q.forward(w)
        assert d.shape == (n_plans, wl.phantom.nvox)
        target, oar = workloads.target_and_oar(wl.phantom)
        m = (target | oar).astype(np.float64)
        P = np.where(target[None, :], 1.2 * d[:, target].mean(axis=1)[:, None], 0.0)
        dLdD = 2.0 * m[None, :] * (d - P)
        g = q.adjoint(dLdD)
        for k in range(n_plans):
            bk = q.plan_beamlets(k)
            jk = [dict(j, plan=0, history_offset=j["first"]) for j in jobs if j["plan"] == k]
            pk = beamlets.BeamletPlan.from_jobs(ctx, jk, 1, src, n_side=20, pitch=0.5, seed=31, n_batches=2)
            gk = pk.adjoint(dLdD[k])
            scale = np.abs(gk).max()
            assert scale > 0, k
            np.testing.assert_allclose(g[bk], gk, rtol=1e-9, atol=1e-12 * scale)
            hit = np.flatnonzero(np.abs(gk) > 1e-2 * scale)
            h = 0.1
            for b in rng.choice(hit, 2, replace=False):
                wk = w[bk]
                wp, wm = wk.copy(), wk.copy()
                wp[b] += h
                wm[b] -= h
                lp = float(np.sum(m * (pk.forward(wp) - P[k]) ** 2))
                lm = float(np.sum(m * (pk.forward(wm) - P[k]) ** 2))
                fd = (lp - lm) / (2 * h)
                assert abs(fd - gk[b]) <= 1e-4 * abs(gk[b]) + 1e-9 * scale, (k, b, fd, gk[b])
    finally:
        ctx.close()
