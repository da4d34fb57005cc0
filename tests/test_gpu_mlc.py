"""Fields that start in the MLC: GPU transport against the oracle.

A synthetic head phase space on the plane z = 16.44 cm (photons and 3%
electrons) is transported through the 60-pair MLC module into the CT
phantom.  The leaf geometry (focused polygons, leaf-end profiles shifted by
the per-field leaf positions), the hand-over from the MLC to the phantom and
the discard of particles that pass outside the leaves are evaluated with the
same binary32 arithmetic on both sides, so per-history energy and step
counts must agree bit for bit, as for fields that start in the phantom."""
import copy

import numpy as np
import pytest

from gegs_amd import head, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ct_mlc():
    return workloads.ct_mlc()


def _load(ctx, wl, src):
    ctx.load_workload(wl)
    ctx.set_source(src)


def test_mlc_plain_per_history_bit_exact(gpu_ctx, ct_mlc):
    src = ct_mlc.source(4000, 3)
    _load(gpu_ctx, ct_mlc, src)
    g = gpu_ctx.run(seed=17, n_batches=2, method="plain", per_history=True)
    o = oracle.run(ct_mlc, src, seed=17, per_history=True)
    assert np.array_equal(g["steps"], o["steps"])
    assert np.array_equal(g["edep"], o["edep"])
    assert g["counters"]["photon_steps"] == o["counters"].photon_steps
    assert g["counters"]["electron_steps"] == o["counters"].electron_steps
    assert np.allclose(g["dose"], o["dose"], rtol=1e-12, atol=1e-12)
    # the aperture matters: part of the histories never reach the phantom
    assert 0.05 < np.mean(o["edep"] == 0) < 0.95


def test_mlc_event_matches_plain(gpu_ctx, ct_mlc):
    src = ct_mlc.source(6000, 4)
    _load(gpu_ctx, ct_mlc, src)
    p = gpu_ctx.run(seed=5, n_batches=2, method="plain", per_history=True)
    e = gpu_ctx.run(seed=5, n_batches=2, method="event", per_history=True)
    assert e["counters"]["overflow"] == 0
    assert np.array_equal(e["steps"], p["steps"])
    assert np.array_equal(e["edep"], p["edep"])
    assert np.allclose(e["dose"], p["dose"], rtol=1e-11, atol=1e-12)


def test_mlc_leaf_sets_per_job(gpu_ctx, ct_mlc):
    """Two fields with different apertures in one batched loop: each job
    follows its own leaf set, history by history."""
    src = ct_mlc.source(5000, 6)
    pos2 = head.random_aperture(11)
    _load(gpu_ctx, ct_mlc, src)
    sid = ct_mlc.field.source_to_isocenter
    f0 = workloads.make_field(True, sid, 0.0, module_first=head.M_MLC)
    f1 = workloads.make_field(True, sid, 72.0, module_first=head.M_MLC)
    both = np.concatenate([ct_mlc.leaf_positions, pos2]).astype(np.float32)
    gpu_ctx.set_leaf_positions(both)
    try:
        jobs = [dict(field=f0, plan=0, first=0, count=3000, leaf_set=0),
                dict(field=f1, plan=1, first=3000, count=2000, leaf_set=1)]
        bt = gpu_ctx.run_batched(jobs, 2, seed=23, n_batches=2, per_history=True)
        assert bt["counters"]["overflow"] == 0
    finally:
        gpu_ctx.set_leaf_positions(ct_mlc.leaf_positions)
    for jb, pos in ((jobs[0], ct_mlc.leaf_positions), (jobs[1], pos2)):
        wl = copy.copy(ct_mlc)
        wl.field = jb["field"]
        wl.leaf_positions = pos
        sl = slice(jb["first"], jb["first"] + 600)
        o = oracle.run(wl, src[sl], seed=23, hist_offset=jb["first"], per_history=True)
        assert np.array_equal(bt["steps"][sl], o["steps"])
        assert np.array_equal(bt["edep"][sl], o["edep"])


def test_jaws_then_mlc_bit_exact(gpu_ctx):
    """Fields starting in the secondary jaws pass Y, X, the MLC and the
    phantom: the plain kernel reproduces the oracle history by history and
    the event path reproduces the plain kernel."""
    wl = workloads.ct_jaws_mlc()
    src = wl.source(4000, 9)
    _load(gpu_ctx, wl, src)
    try:
        g = gpu_ctx.run(seed=31, n_batches=2, method="plain", per_history=True)
        e = gpu_ctx.run(seed=31, n_batches=2, method="event", per_history=True)
    finally:
        gpu_ctx.set_field(workloads.ct_mlc().field)
    o = oracle.run(wl, src, seed=31, per_history=True)
    assert np.array_equal(g["steps"], o["steps"])
    assert np.array_equal(g["edep"], o["edep"])
    assert np.array_equal(e["steps"], g["steps"])
    assert np.allclose(e["edep"], g["edep"], rtol=1e-12, atol=1e-12)
    assert e["counters"]["overflow"] == 0
