"""gegs_run_plain_noalloc against gegs_run_plain, through the C ABI.

The no-alloc entry point runs the same plain kernel with its field, head
geometry, batch row, work counters and particle stacks in the library's fixed
device array instead of context allocations, and writes the caller's device
dose buffers.  Histories do not depend on the grid, so the counters must be
identical and the dose may differ only by the order of the FP64 tally
additions.  Two calls in a row on one context, the second with fewer
histories and a smaller phantom region of the source, show that the array's
previous contents do not leak into the next call."""
import numpy as np
import pytest

from gegs_amd import workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _device_run(ctx, seed, n_batches, count, grid=0):
    nvox = ctx.nvox
    dose = torch.empty(nvox, dtype=torch.float64, device="cuda")
    dose2 = torch.empty(nvox, dtype=torch.float64, device="cuda")
    # the library runs on its own stream: torch's work on these buffers must be done
    torch.cuda.synchronize()
    cnt = ctx.run_plain_noalloc(seed, dose.data_ptr(), dose2.data_ptr(), n_batches=n_batches, count=count,
                                grid=grid)
    return dose.cpu().numpy(), dose2.cpu().numpy(), cnt


def _check(ref, dose, dose2, cnt):
    assert cnt["overflow"] == 0
    for k in ("photon_steps", "electron_steps", "region_gathers", "deposits", "histories", "interactions"):
        assert cnt[k] == ref["counters"][k], k
    assert np.allclose(dose, ref["dose"], rtol=1e-12, atol=1e-12)
    assert np.allclose(dose2, ref["dose2"], rtol=1e-12, atol=1e-14)


def test_noalloc_matches_plain_twice(gpu_ctx):
    wl = workloads.ct_mlc()
    src = wl.source(6000, 3)
    gpu_ctx.load_workload(wl)
    gpu_ctx.set_source(src)
    print("no-alloc: ct_mlc, 6000 then 2500 records, seed 17", flush=True)
    ref1 = gpu_ctx.run(seed=17, n_batches=3, method="plain", count=6000)
    d1, q1, c1 = _device_run(gpu_ctx, 17, 3, 6000)
    _check(ref1, d1, q1, c1)
    assert c1["histories"] == 6000
    ref2 = gpu_ctx.run(seed=29, n_batches=2, method="plain", count=2500)
    d2, q2, c2 = _device_run(gpu_ctx, 29, 2, 2500)
    _check(ref2, d2, q2, c2)
    assert c2["histories"] == 2500


def test_noalloc_rejects_a_grid_beyond_the_scratch(gpu_ctx):
    from gegs_amd import native
    wl = workloads.water()
    gpu_ctx.load_workload(wl)
    gpu_ctx.set_source(wl.source(500, 1))
    # 128 threads x 40 slots x 64 B per block: 100000 blocks need 33 GB
    with pytest.raises(native.GegsError, match="scratch"):
        _device_run(gpu_ctx, 1, 1, 500, grid=100000)
    d, q, c = _device_run(gpu_ctx, 1, 1, 500)
    ref = gpu_ctx.run(seed=1, n_batches=1, method="plain", count=500)
    _check(ref, d, q, c)
