"""Fields that start in the block or the wedge: GPU transport against the
oracle.

The synthetic head phase space on the plane z = 16.44 cm passes a tungsten
block (z 55-58 cm) and/or a tungsten wedge (z 60-61.5 cm), then the CT
phantom.  The prism tests, the module hand-over (block -> wedge -> phantom)
and the step of a particle that misses a prism are evaluated with the same
binary32 arithmetic on both sides, so per-history energy and step counts
agree bit for bit; the event path reproduces the plain kernel."""
import numpy as np
import pytest

from gegs_amd import head, native, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("block,wedge", [(True, True), (False, True), (True, False)])
def test_block_wedge_bit_exact(block, wedge):
    wl = workloads.ct_block_wedge(block=block, wedge=wedge)
    src = wl.source(20000, 5)
    print("block_wedge block=%s wedge=%s seed 37, 20000 histories" % (block, wedge), flush=True)
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        g = ctx.run(seed=37, n_batches=2, method="plain", per_history=True)
        e = ctx.run(seed=37, n_batches=2, method="event", per_history=True)
    finally:
        ctx.close()
    o = oracle.run(wl, src, seed=37, per_history=True, threads=4)
    assert np.array_equal(g["steps"], o["steps"])
    assert np.array_equal(g["edep"], o["edep"])
    assert g["counters"]["photon_steps"] == o["counters"].photon_steps
    assert np.allclose(g["dose"], o["dose"], rtol=1e-12, atol=1e-12)
    assert e["counters"]["overflow"] == 0
    assert np.array_equal(e["steps"], g["steps"])
    assert np.allclose(e["edep"], g["edep"], rtol=1e-12, atol=1e-12)
    # part of the histories cross the modules and reach the phantom (through 3 cm of tungsten
    # and the wedge's 1.5 cm slab, about 0.6%; through the block alone, about 1.4%; through
    # the wedge alone, about 14%)
    assert 0.003 < np.mean(o["edep"] > 0) < 0.5


def test_block_field_needs_a_block():
    wl = workloads.ct_block_wedge(block=True, wedge=False)
    src = wl.source(100, 1)
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        ctx.load_block_wedge(head.build_block_wedge(None, workloads.WEDGE_PYRAMIDS))
        with pytest.raises(native.GegsError, match="no block is loaded"):
            ctx.run(seed=1, n_batches=1, method="plain")
        bad = head.build_block_wedge(None, workloads.WEDGE_PYRAMIDS)
        bad.wedge_nver = 7
        with pytest.raises(native.GegsError, match="3 to 6 vertices"):
            ctx.load_block_wedge(bad)
    finally:
        ctx.close()
