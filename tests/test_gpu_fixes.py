"""Opt-in correction of quirk Q1 (GEGS_FIX_Q1: hownear_phantom measures from
the particle's own voxel, where the reference takes the next one,
geom.cu:1487-1490).

With the fix on, the GPU still reproduces the oracle history by history
(both sides switch the same decode), the event path still reproduces the
plain kernel, electron steps become longer (fewer steps: the step limit is
the true distance to the voxel boundary, no longer a distance <= 0 that
forces the minimum step), and the dose stays statistically equivalent to
the reference's (per-voxel z-scores of independent runs)."""
import numpy as np
import pytest

from gegs_amd import native, stats, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ct_small():
    return workloads.ct()


def test_q1_fix_bit_exact_and_fewer_electron_steps(ct_small):
    src = ct_small.source(20000, 21)
    print("Q1 fix: synthetic CT workload, 20000 histories, seed 41", flush=True)
    ctx = native.Context(0)
    try:
        ctx.load_workload(ct_small)
        ctx.set_source(src)
        q = ctx.run(seed=41, n_batches=2, method="plain", per_history=True)
        ctx.set_fixes(native.FIX_Q1)
        g = ctx.run(seed=41, n_batches=2, method="plain", per_history=True)
        e = ctx.run(seed=41, n_batches=2, method="event", per_history=True)
    finally:
        ctx.close()
    oracle.set_fixes(native.FIX_Q1)
    try:
        o = oracle.run(ct_small, src, seed=41, per_history=True, threads=8)
    finally:
        oracle.set_fixes(0)
    assert np.array_equal(g["steps"], o["steps"])
    assert np.array_equal(g["edep"], o["edep"])
    assert g["counters"]["electron_steps"] == o["counters"].electron_steps
    assert e["counters"]["overflow"] == 0
    assert np.array_equal(e["steps"], g["steps"])
    assert np.allclose(e["edep"], g["edep"], rtol=1e-12, atol=1e-12)
    print("electron steps: reference %d, Q1 fix %d" % (q["counters"]["electron_steps"],
                                                      g["counters"]["electron_steps"]), flush=True)
    assert g["counters"]["electron_steps"] < 0.9 * q["counters"]["electron_steps"]


def test_q1_fix_dose_equivalent_to_reference(ct_small):
    n, nb = 2_000_000, 20
    ctx = native.Context(0)
    try:
        ctx.load_workload(ct_small)
        ctx.set_source(ct_small.source(n, 31))
        a = ctx.run(seed=3, n_batches=nb, method="event")
        ctx.set_source(ct_small.source(n, 32))
        ctx.set_fixes(native.FIX_Q1)
        b = ctx.run(seed=4, n_batches=nb, method="event", history_offset=n)
    finally:
        ctx.close()
    sa = stats.batch_sigma(a["dose"], a["dose2"], nb)
    sb = stats.batch_sigma(b["dose"], b["dose2"], nb)
    mean = 0.5 * (a["dose"] + b["dose"])
    mask = mean > 0.3 * mean.max()
    chi2, meanz, nv = stats.zscore_chi2(a["dose"], sa, b["dose"], sb, mask)
    print("Q1 fix vs reference: chi2/voxel %.3f, mean z %.3f over %d voxels" % (chi2, meanz, nv), flush=True)
    assert nv > 200
    assert 0.75 < chi2 < 1.35, (chi2, meanz, nv)
    assert abs(meanz) < 0.2


@pytest.mark.parametrize("mask,gantry", [(native.FIX_Q1 | native.FIX_Q8 | native.FIX_Q7, 0.0),
                                         (native.FIX_ALL, 72.0), (native.FIX_FP64_FRAME, 216.0)])
def test_all_fixes_jaws_mlc_bit_exact(mask, gantry):
    """GEGS_FIX_Q1 | GEGS_FIX_Q8 | GEGS_FIX_Q7, every fix with the binary64
    frame transform, and that transform alone, on the jaws -> MLC -> CT
    workload (range rejection is on for the MLC medium) at three gantry
    angles: the GPU reproduces the oracle with the same mask history by
    history, and the event path reproduces the plain kernel."""
    wl = workloads.ct_jaws_mlc(gantry_deg=gantry)
    src = wl.source(4000, 19)
    print("fixes mask %d, gantry %.0f: jaws + MLC workload, 4000 histories, seed 43" % (mask, gantry), flush=True)
    ctx = native.Context(0)
    try:
        ctx.load_workload(wl)
        ctx.set_source(src)
        ctx.set_fixes(mask)
        g = ctx.run(seed=43, n_batches=2, method="plain", per_history=True)
        e = ctx.run(seed=43, n_batches=2, method="event", per_history=True)
    finally:
        ctx.close()
    oracle.set_fixes(mask)
    try:
        o = oracle.run(wl, src, seed=43, per_history=True, threads=8)
    finally:
        oracle.set_fixes(0)
    assert np.array_equal(g["steps"], o["steps"])
    assert np.array_equal(g["edep"], o["edep"])
    assert e["counters"]["overflow"] == 0
    assert np.array_equal(e["steps"], g["steps"])
    assert np.allclose(e["edep"], g["edep"], rtol=1e-12, atol=1e-12)
