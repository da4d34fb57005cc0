"""Edge cases of the C ABI on the GPU: empty and out-of-range inputs are
status codes before any launch; a single history, ragged batches (histories
not divisible by the batch count) and the largest batch count give the same
histories as the plain kernel and the oracle.

The reference has no such checks (a short phase-space read returns -1 and
the batch is skipped, CUDA_EGS.cu:462-476); here every refusal names its
argument, and the transported histories do not depend on how they are
batched."""
import numpy as np
import pytest

from gegs_amd import native, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ct_small():
    return workloads.ct()


def test_empty_and_out_of_range_inputs_are_refused(gpu_ctx, ct_small):
    ctx = gpu_ctx
    ctx.load_workload(ct_small)
    ctx.set_source(ct_small.source(0, 1))                     # an empty source is accepted ...
    for method in ("plain", "event"):
        with pytest.raises(native.GegsError) as e:            # ... and a run over it is a call-order error
            ctx.run(seed=1, n_batches=2, method=method)
        assert e.value.code == 90002
    src = ct_small.source(100, 2)
    ctx.set_source(src)
    f = ct_small.field
    bad = [dict(method="plain", first=100), dict(method="plain", first=0, count=101),
           dict(method="event", first=100), dict(method="event", first=50, count=51),
           dict(method="plain", n_batches=0), dict(method="event", n_batches=4097)]
    for kw in bad:
        args = dict(seed=1, n_batches=2)
        args.update(kw)
        with pytest.raises(native.GegsError) as e:
            ctx.run(**args)
        assert e.value.code == 2001, kw
    for jobs in ([], [dict(field=f, plan=0, first=0, count=0)], [dict(field=f, plan=0, first=90, count=11)],
                 [dict(field=f, plan=1, first=0, count=10)]):
        with pytest.raises(native.GegsError) as e:
            ctx.run_batched(jobs, 1, seed=1, n_batches=2)
        assert e.value.code == 2001, jobs
    # the context is still usable after the refusals
    r = ctx.run(seed=1, n_batches=2, method="event")
    assert r["counters"]["histories"] == 100 and r["dose"].sum() > 0


@pytest.fixture(scope="module")
def water16():
    return workloads.water(n=16, voxel=2.0)


@pytest.mark.parametrize("wl_name,count,n_batches", [("ct", 1, 1), ("ct", 1001, 7), ("water16", 257, 4096)])
def test_single_history_ragged_and_maximal_batches(gpu_ctx, ct_small, water16, wl_name, count, n_batches):
    """Per-history energy and steps of the event loop equal the plain
    kernel's and the oracle's whatever the batching (a batch of 1001
    histories in 7 slices, 257 histories in 4096 batches, most of them
    empty; the event loop keeps one dose row per batch, so the largest batch
    count runs on a 16^3 phantom), and the batch doses add up to the run's
    dose."""
    print("edges: %s, %d histories in %d batches" % (wl_name, count, n_batches), flush=True)
    wl = ct_small if wl_name == "ct" else water16
    ctx = gpu_ctx
    ctx.load_workload(wl)
    src = wl.source(count, 77)
    ctx.set_source(src)
    p = ctx.run(seed=5, n_batches=n_batches, method="plain", per_history=True)
    e = ctx.run(seed=5, n_batches=n_batches, method="event", per_history=True)
    o = oracle.run(wl, src, seed=5, per_history=True)
    assert np.array_equal(p["steps"], o["steps"]) and np.array_equal(p["edep"], o["edep"])
    assert np.array_equal(e["steps"], o["steps"])
    assert np.allclose(e["edep"], o["edep"], rtol=1e-12, atol=1e-12)
    assert e["counters"]["histories"] == count and not e["counters"]["overflow"]
    one = ctx.run(seed=5, n_batches=1, method="event")
    assert np.allclose(e["dose"], one["dose"], rtol=1e-9, atol=1e-12 * max(1.0, one["dose"].max()))


def test_sweep_streams_give_the_same_histories(gpu_ctx, ct_small):
    """The photon and electron sides of a sweep on two streams (the default)
    or one after the other on one stream: the same histories (energy and
    steps per history equal bit for bit), the same dose to the order of the
    FP64 sums; other stream counts are refused."""
    print("edges: sweep_streams 1 and 2", flush=True)
    ctx = gpu_ctx
    ctx.load_workload(ct_small)
    src = ct_small.source(3000, 12)
    ctx.set_source(src)
    base = ctx.get_tuning()
    assert base["sweep_streams"] == 2
    runs = {}
    try:
        for n in (1, 2):
            ctx.set_tuning(dict(base, sweep_streams=n))
            assert ctx.get_tuning()["sweep_streams"] == n
            runs[n] = ctx.run(seed=3, n_batches=3, method="event", per_history=True)
        with pytest.raises(native.GegsError) as e:
            ctx.set_tuning(dict(base, sweep_streams=3))
        assert e.value.code == 2001
    finally:
        ctx.set_tuning(base)
    assert np.array_equal(runs[1]["steps"], runs[2]["steps"])
    assert np.allclose(runs[1]["edep"], runs[2]["edep"], rtol=1e-12, atol=1e-12)
    d1, d2 = runs[1]["dose"], runs[2]["dose"]
    assert d2.sum() > 0 and np.allclose(d1, d2, rtol=1e-9, atol=1e-12 * d2.max())


@pytest.mark.parametrize("tail", [0xFFFFFFFF, 64, 1 << 18])
def test_tail_launch_gives_the_same_histories(gpu_ctx, ct_small, tail):
    """The run's last particles finished by one launch (each with its
    descendants in one thread, as the plain kernel transports them) or by
    the sweeps to the end: the same histories as the oracle, bit for bit in
    steps, whichever sweep hands them over (a tail of 64 particles starts
    late in the run, 2^18 right after the first replay of this small run)."""
    print("edges: tail_particles %d" % tail, flush=True)
    ctx = gpu_ctx
    ctx.load_workload(ct_small)
    src = ct_small.source(4000, 21)
    ctx.set_source(src)
    base = ctx.get_tuning()
    try:
        ctx.set_tuning(dict(base, tail_particles=tail, photon_chunk=16, electron_chunk=8))
        assert ctx.get_tuning()["tail_particles"] == tail
        r = ctx.run(seed=9, n_batches=2, method="event", per_history=True)
        with pytest.raises(native.GegsError) as e:
            ctx.set_tuning(dict(base, tail_particles=(1 << 18) + 1))
        assert e.value.code == 2001
    finally:
        ctx.set_tuning(base)
    o = oracle.run(ct_small, src, seed=9, per_history=True)
    assert np.array_equal(r["steps"], o["steps"])
    assert np.allclose(r["edep"], o["edep"], rtol=1e-12, atol=1e-12)
    assert r["counters"]["histories"] == 4000 and not r["counters"]["overflow"]
