"""GPU transport against the oracle, through the C ABI.

With the same seed both sides draw the same Philox streams and evaluate the
same binary32 arithmetic (no FMA contraction, correctly rounded division and
square root, the same elementary-function algorithms), so trajectories agree
history by history: per-history energy and step counts must be identical, and
per-voxel dose may differ only by the order of the FP64 tally additions."""
import numpy as np
import pytest

from gegs_amd import native, stats, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water():
    return workloads.water()


@pytest.fixture(scope="module")
def ct_small():
    return workloads.ct()


def _run_gpu(ctx, wl, src, seed, method="plain", per_history=True, n_batches=4):
    ctx.load_workload(wl)
    ctx.set_source(src)
    return ctx.run(seed=seed, n_batches=n_batches, method=method, per_history=per_history)


def test_water_per_history_bit_exact(gpu_ctx, water):
    src = water.source(4000, 1)
    g = _run_gpu(gpu_ctx, water, src, 1234)
    o = oracle.run(water, src, seed=1234, per_history=True)
    assert np.array_equal(g["edep"], o["edep"])
    assert np.array_equal(g["steps"], o["steps"])
    c = o["counters"]
    gc = g["counters"]
    assert gc["photon_steps"] == c.photon_steps
    assert gc["electron_steps"] == c.electron_steps
    assert gc["deposits"] == c.deposits
    assert gc["interactions"] == list(c.interactions)
    assert np.allclose(g["dose"], o["dose"], rtol=1e-12, atol=1e-12)


def test_ct_per_history_bit_exact(gpu_ctx, ct_small):
    src = ct_small.source(3000, 2)
    g = _run_gpu(gpu_ctx, ct_small, src, 99)
    o = oracle.run(ct_small, src, seed=99, per_history=True)
    assert np.array_equal(g["edep"], o["edep"])
    assert np.array_equal(g["steps"], o["steps"])
    assert np.allclose(g["dose"], o["dose"], rtol=1e-12, atol=1e-12)


EVENTS = [(0, 0, (0.02, 6.0)), (1, 0, (0.02, 6.0)), (2, 0, (0.012, 0.5)), (3, 0, (0.011, 1.0)),
          (4, 0, (1.03, 6.0)), (5, -1, (0.72, 6.5)), (5, 1, (0.72, 6.5)), (6, -1, (1.0, 6.5)),
          (7, -1, (0.75, 6.5)), (8, 1, (0.75, 6.5)), (9, 1, (0.75, 6.5))]


@pytest.mark.parametrize("status,charge,erange", EVENTS)
def test_samplers_uniform_for_uniform(gpu_ctx, ct_small, status, charge, erange):
    """Every event handler, fed the same stream, returns bit-identical
    particles and scored energy on the GPU and in the oracle."""
    _samplers_agree(gpu_ctx, ct_small, status, charge, erange)


@pytest.mark.parametrize("mask,status,charge,erange", [(0, 4, 0, (50.5, 80.0)), (native.FIX_Q2, 4, 0, (50.5, 80.0)),
                                                       (native.FIX_Q3, 1, 0, (0.02, 6.0))])
def test_quirk_switch_samplers_uniform_for_uniform(gpu_ctx, ct_small, mask, status, charge, erange):
    """The pair sampler above 50 MeV (the Coulomb-corrected branch, where
    quirk Q2 reads bpar) with and without GEGS_FIX_Q2, and Compton with
    GEGS_FIX_Q3: GPU and oracle agree uniform for uniform under each mask."""
    gpu_ctx.set_fixes(mask)
    oracle.set_fixes(mask)
    try:
        _samplers_agree(gpu_ctx, ct_small, status, charge, erange)
    finally:
        gpu_ctx.set_fixes(0)
        oracle.set_fixes(0)


def _samplers_agree(gpu_ctx, ct_small, status, charge, erange):
    gpu_ctx.load_workload(ct_small)
    ph = ct_small.phantom
    rng = np.random.default_rng(status * 10 + charge + 3)
    n = 300
    parts = np.zeros(n, dtype=native.PARTICLE_DTYPE)
    ix = rng.integers(40, 88, n)
    iy = rng.integers(40, 88, n)
    iz = rng.integers(30, 70, n)
    region = 6 + ix + iy * ph.nx + iz * ph.nx * ph.ny
    med = ct_small.regions["med"][region]
    keep = med != 0xFFFF
    # head-frame coordinates of the voxel centres (CT axis convention, gantry 0)
    xc = 0.5 * (ph.xb[ix] + ph.xb[ix + 1])
    yc = 0.5 * (ph.yb[iy] + ph.yb[iy + 1])
    zc = 0.5 * (ph.zb[iz] + ph.zb[iz + 1])
    sid = ct_small.field.source_to_isocenter
    parts["x"], parts["y"], parts["z"] = zc, xc, yc + np.float32(sid)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    parts["u"], parts["v"], parts["w"] = d[:, 0], d[:, 1], d[:, 2]
    parts["e"] = rng.uniform(*erange, n)
    parts["wt"] = 1.0
    parts["dnear"] = 1e5
    parts["region"] = region
    parts["charge"] = charge
    parts["status"] = status
    parts["lin"] = rng.integers(1, 2 ** 62, n, dtype=np.uint64)
    parts["blk"] = rng.integers(0, 50, n)
    parts = parts[keep]
    out, nout, edep = gpu_ctx.sample_events(parts, seed=321, max_out=24)
    for i, p in enumerate(parts):
        o, e = oracle.interaction(ct_small, p, 321)
        assert nout[i] == len(o), i
        assert e == edep[i], i
        assert out[i, :len(o)].tobytes() == o.tobytes(), i


def test_dose_statistically_equivalent_across_seeds(gpu_ctx, ct_small):
    """Independent seeds: per-voxel z-scores of GPU vs oracle dose follow a
    unit normal (chi^2 per voxel near 1) over voxels above 20% of Dmax."""
    n = 200000
    g = _run_gpu(gpu_ctx, ct_small, ct_small.source(n, 8), 1, per_history=False, n_batches=20)
    o = oracle.run(ct_small, ct_small.source(n, 9), seed=2, threads=8)
    sg = stats.batch_sigma(g["dose"], g["dose2"], 20)
    so = stats.history_sigma(o["dose"], o["dose2"], n)
    # select on the mean of both runs so that neither side's noise biases z
    mean = 0.5 * (g["dose"] + o["dose"])
    mask = mean > 0.3 * mean.max()
    chi2, meanz, nv = stats.zscore_chi2(g["dose"], sg, o["dose"], so, mask)
    assert nv > 200
    assert 0.75 < chi2 < 1.35, (chi2, meanz, nv)
    assert abs(meanz) < 0.2


def test_errors_are_status_codes(gpu_ctx):
    fresh = native.Context(0)
    with pytest.raises(native.GegsError) as e:
        fresh.run(seed=1)
    assert e.value.code == 90002
    fresh.close()


TUNINGS = {
    "default": None,
    # every photon crossing and electron step is suspended and resumed in the
    # next sweep, and primaries enter a few thousand at a time
    "suspend_every_step": dict(photon_chunk=1, electron_chunk=1, inject=4096, threshold=8192,
                               block=128, sweeps_per_graph=16),
}


@pytest.mark.parametrize("wl_name,tuning", [("water", "default"), ("ct", "default"),
                                            ("ct", "suspend_every_step")])
def test_event_path_reproduces_histories(gpu_ctx, wl_name, tuning, water, ct_small):
    """The event-queued path transports every history along the same
    trajectory as the plain kernel and the oracle: identical step counts and
    event counters; per-history energy equal up to the order of the FP64
    additions (deposits are summed per particle, then per history)."""
    wl = water if wl_name == "water" else ct_small
    n = 20000 if tuning == "default" else 4000
    src = wl.source(n, 31)
    gpu_ctx.load_workload(wl)
    gpu_ctx.set_source(src)
    gpu_ctx.set_tuning(TUNINGS[tuning])
    try:
        ev = gpu_ctx.run(seed=5, n_batches=3, method="event", per_history=True)
    finally:
        gpu_ctx.set_tuning(None)
    pl = gpu_ctx.run(seed=5, n_batches=3, method="plain", per_history=True)
    assert np.array_equal(ev["steps"], pl["steps"])
    assert np.allclose(ev["edep"], pl["edep"], rtol=1e-12, atol=1e-12)
    for k in ("photon_steps", "electron_steps", "region_gathers", "deposits", "histories"):
        assert ev["counters"][k] == pl["counters"][k], k
    assert ev["counters"]["interactions"] == pl["counters"]["interactions"]
    assert ev["counters"]["overflow"] == 0
    assert np.allclose(ev["dose"], pl["dose"], rtol=1e-12, atol=1e-12)
    o = oracle.run(wl, src[:2000], seed=5, per_history=True)
    assert np.array_equal(ev["steps"][:2000], o["steps"])


def test_tuning_roundtrip_and_validation(gpu_ctx):
    gpu_ctx.set_tuning(dict(photon_chunk=7, inject=1000))
    t = gpu_ctx.get_tuning()
    assert t["photon_chunk"] == 7 and t["inject"] == 1000 and t["block"] == 256
    with pytest.raises(native.GegsError) as e:
        gpu_ctx.set_tuning(dict(block=96))
    assert e.value.code == 2001
    # a bank is 32 append segments of at least one slot each
    with pytest.raises(native.GegsError) as e:
        gpu_ctx.set_tuning(dict(capacity=31))
    assert e.value.code == 2001
    gpu_ctx.set_tuning(None)
    assert gpu_ctx.get_tuning()["photon_chunk"] == 32


def test_batched_fields_and_plans_match_per_field_runs(gpu_ctx, ct_small):
    """Three fields of two plans in one event loop give, history by history,
    the trajectories of separate per-field runs of the plain kernel, and each
    plan's dose is the sum of its fields' doses."""
    import copy
    from gegs_amd import workloads
    wl = ct_small
    src = wl.source(6000, 41)
    gpu_ctx.load_workload(wl)
    gpu_ctx.set_source(src)
    sid = wl.field.source_to_isocenter
    fields = [workloads.make_field(True, sid, g) for g in (0.0, 90.0, 200.0)]
    jobs = [dict(field=fields[0], plan=0, first=0, count=3000),
            dict(field=fields[1], plan=1, first=3000, count=2000),
            dict(field=fields[2], plan=0, first=5000, count=1000)]
    bt = gpu_ctx.run_batched(jobs, 2, seed=11, n_batches=3, per_history=True)
    assert bt["counters"]["overflow"] == 0
    ref_dose = np.zeros((2, gpu_ctx.nvox))
    slot = 0
    try:
        for jb in jobs:
            gpu_ctx.set_field(jb["field"])
            pl = gpu_ctx.run(seed=11, n_batches=2, method="plain", per_history=True,
                             first=jb["first"], count=jb["count"])
            sl = slice(jb["first"], jb["first"] + jb["count"])
            assert np.array_equal(bt["steps"][slot:slot + jb["count"]], pl["steps"][sl])
            assert np.allclose(bt["edep"][slot:slot + jb["count"]], pl["edep"][sl], rtol=1e-12, atol=1e-12)
            ref_dose[jb["plan"]] += pl["dose"]
            slot += jb["count"]
    finally:
        gpu_ctx.set_field(wl.field)
    assert np.allclose(bt["dose"], ref_dose, rtol=1e-11, atol=1e-12)
    # the rotated field against the oracle
    wl90 = copy.copy(wl)
    wl90.field = fields[1]
    o = oracle.run(wl90, src[3000:3500], seed=11, hist_offset=3000, per_history=True)
    assert np.array_equal(bt["steps"][3000:3500], o["steps"])


def test_kernel_timing_leaves_results_unchanged(gpu_ctx, ct_small):
    """Timing the photon kernel with event nodes inside the replayed graphs
    changes nothing in the transport and reports one duration per launch."""
    src = ct_small.source(20000, 8)
    gpu_ctx.load_workload(ct_small)
    gpu_ctx.set_source(src)
    ref = gpu_ctx.run(seed=3, n_batches=1, method="event", per_history=True)
    gpu_ctx.set_kernel_timing(True)
    try:
        t = gpu_ctx.run(seed=3, n_batches=1, method="event", per_history=True)
        ms, launches = gpu_ctx.kernel_timing()
    finally:
        gpu_ctx.set_kernel_timing(False)
    assert np.array_equal(t["steps"], ref["steps"])
    assert np.array_equal(t["edep"], ref["edep"])
    sweeps_per_graph = gpu_ctx.get_tuning()["sweeps_per_graph"]
    assert launches > 0 and launches % sweeps_per_graph == 0
    # the photon kernel is part of the run bracketed by the run's own events
    assert 0.0 < ms <= 1.02 * t["ms"] + 0.5


def test_atomic_probe_rate_and_bounds(gpu_ctx):
    idx = np.random.default_rng(3).integers(0, 1000, 1 << 16).astype(np.uint32)
    assert gpu_ctx.atomic_probe(idx, 1000, reps=2) > 0
    bad = idx.copy()
    bad[7] = 1000
    with pytest.raises(native.GegsError) as e:
        gpu_ctx.atomic_probe(bad, 1000)
    assert e.value.code == 2001


def test_media_with_nbr_split_below_one_are_refused(gpu_ctx, ct_small):
    """nbr_split < 1 would drop in-flight annihilations with their energy in
    the reference (quirk Q5): gegs_load_media refuses such tables."""
    t = ct_small.tables
    old = t.c.nbr_split
    t.c.nbr_split = 0
    try:
        with pytest.raises(native.GegsError, match="nbr_split"):
            gpu_ctx.load_media(t)
    finally:
        t.c.nbr_split = old
    gpu_ctx.load_media(t)
