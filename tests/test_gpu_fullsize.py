"""Full-size workloads on the GPU, checked through properties that hold at any
size, where the oracle would take too long: the event-queued loop (one
field, or all fields of a plan in one loop) and the plain per-history kernel
transport the same histories with the same Philox streams, so every integer
counter (photon and electron steps, region gathers, deposits, pops per event
type) is identical, the doses agree up to the order of the FP64 additions,
and no particle is lost to a full bank.

Sizes: the 128x128x100 CT at 2.5 mm with 1e7 histories in 10 batches, and the
5-field MLC plan on the same CT with 2e7 phase-space records (plain: five
per-field runs; batched: one run of the five fields)."""
import numpy as np
import pytest

from gegs_amd import native, workloads

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
