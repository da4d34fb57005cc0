"""The event-method autotuner on a small sample of the CT workload: it walks
its search space with the production kernels, leaves the best tuning set on
the context, and the tuning it picks changes speed only -- the event loop
under it reproduces the plain kernel history by history."""
import numpy as np
import pytest

from gegs_amd import autotune, workloads

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_valid_tuning_that_keeps_results(gpu_ctx):
    wl = workloads.ct()
    src = wl.source(40000, 12)
    gpu_ctx.load_workload(wl)
    gpu_ctx.set_source(src)
    try:
        tuning, best, trials = autotune.autotune(gpu_ctx, 0, 20000, seed=3, repeat=1)
        assert best > 0 and len(trials) > len(autotune.SPACE)
        for k, v in tuning.items():
            assert k in autotune.START
            if k in autotune.SPACE:
                assert v in autotune.SPACE[k] or v == autotune.START[k]
        got = gpu_ctx.get_tuning()
        for k in ("photon_chunk", "electron_chunk", "refill_min", "refill_slots", "flight_block"):
            assert got[k] == tuning[k]
        e = gpu_ctx.run(seed=8, n_batches=2, method="event", first=20000, count=20000, per_history=True)
        p = gpu_ctx.run(seed=8, n_batches=2, method="plain", first=20000, count=20000, per_history=True)
    finally:
        gpu_ctx.set_tuning(None)
    sl = slice(20000, 40000)
    assert e["counters"]["overflow"] == 0
    assert np.array_equal(e["steps"][sl], p["steps"][sl])
    assert np.allclose(e["edep"][sl], p["edep"][sl], rtol=1e-12, atol=1e-12)
