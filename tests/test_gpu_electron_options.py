"""The electron-step options the reference selects at compile time --
PRESTA-II (ESA_PII, msdist_pII, electron.cu:282-614) and the exact boundary
crossing with single scattering (BCA_EXACT, electron.cu:1718-1752,
:2247-2276) -- chosen here by the loaded tables: the GPU (plain kernel and
event loop, which then run their general instantiations) reproduces the
oracle history by history on the CT workload."""
import numpy as np
import pytest

from gegs_amd import hatch, native, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("transport,bca", [("presta-II", "inexact"), ("presta-I", "exact"), ("presta-II", "exact")])
def test_electron_options_bit_exact(transport, bca):
    wl = workloads.ct()
    wl._tables = hatch.MediaTables(wl.media_names, mlc_medium=wl.mlc_medium, transport=transport, bca=bca)
    src = wl.source(3000, 41)
    print("electron options %s / %s BCA: CT, 3000 histories, seed 47" % (transport, bca), flush=True)
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        g = ctx.run(seed=47, n_batches=2, method="plain", per_history=True)
        e = ctx.run(seed=47, n_batches=2, method="event", per_history=True)
    finally:
        ctx.close()
    o = oracle.run(wl, src, seed=47, per_history=True, threads=8)
    assert np.array_equal(g["steps"], o["steps"])
    assert np.array_equal(g["edep"], o["edep"])
    assert g["counters"]["electron_steps"] == o["counters"].electron_steps
    assert np.allclose(g["dose"], o["dose"], rtol=1e-12, atol=1e-12)
    assert e["counters"]["overflow"] == 0
    assert np.array_equal(e["steps"], g["steps"])
    assert np.allclose(e["edep"], g["edep"], rtol=1e-12, atol=1e-12)


def test_spin_tables_are_refused(gpu_ctx):
    """Spin effects need the reference's spin data (absent); tables that ask
    for them are refused instead of run without."""
    wl = workloads.water(n=12, voxel=2.5)
    t = hatch.MediaTables(wl.media_names)
    t.c.spin_effects = 1
    with pytest.raises(native.GegsError, match="spin"):
        gpu_ctx.load_media(t)
