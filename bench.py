#!/usr/bin/env python3
"""Benchmark of the coupled electron-photon transport on one MI355X.

Workload: the heterogeneous-CT case -- seeded synthetic CT of 128x128x100
voxels at 2.5 mm (air / lung / ICRU tissue / ICRP bone through the ctcreate
ramp), point source at SSD 100 cm, 10x10 cm^2 field, 24-bin 6 MV-like
spectrum; 1e7 histories by default (10 steps of 1e6).  Beside it, the
5-field MLC plan (gantry 0/72/144/216/288 degrees, seeded apertures, uneven
field weights, 2e7 phase-space records with 3% electrons) as five plain
launches per step against one batched event loop per step.

A step is one statistical batch of 1e6 source records already resident in
HBM; the timed call transports all steps, once with the plain per-history
kernel (the baseline method, one launch per batch) and once with the batched
event-queued loop (the headline, all batches in one loop).  Reported:
histories/s of the whole job, dose efficiency 1/(sigma^2 T) (sigma = mean
relative uncertainty over voxels above 50% of Dmax, from the batch
variance), the HBM
roofline of the event loop's dominant kernel (the photon flight kernel,
timed live with HIP events), the adjoint gradient rate of a 5-field,
2000-beamlet plan on the 128^3 CT, a queue of 16 plans (93 fields, 1e8
histories) on a 256x256x180 CT batched vs field by field, and the CPU
oracle timed on this host's cores on a bounded sample.

  python bench.py [--gpus 1] [--steps 10] [--warmup 1] [--histories-per-step 1000000]
                  [--mlc-plan-histories 20000000] [--adjoint-histories 10000000]
                  [--plan-queue-histories 100000000] [--tuning JSON]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SOURCE_BYTES = 36          # one gegs_source record
REGION_BYTES = 16          # one region record gathered on voxel entry
TABLE_ROW_BYTES = 8        # one float2 table row
DEPOSIT_BYTES = 16         # FP64 read-modify-write
# table rows read per transport step (photon: ge, gmfp, cohe; electron: eke01,
# sig, dedx x2, tmxs, e_array/range rows)
ROWS_PER_PHOTON_STEP = 3
ROWS_PER_ELECTRON_STEP = 6
# photon flight kernel: the region record of the voxel and the three photon
# rows, per photon step
PHOTON_STEP_BYTES = REGION_BYTES + ROWS_PER_PHOTON_STEP * TABLE_ROW_BYTES


def profile_figure(key):
    """A figure of the committed rocprofv3 summary of the photon kernel
    (profiles/photon_kernel_ct128_event.md, a line `key: value`), or None:
    hbm_bytes_per_photon_step (FETCH_SIZE / WRITE_SIZE, copy-calibrated) and
    valu_issue_frac (SQ_ACTIVE_INST_VALU over the SIMDs' cycles)."""
    import re
    path = os.path.join(ROOT, "profiles", "photon_kernel_ct128_event.md")
    try:
        with open(path) as f:
            m = re.search(r"^%s: ([0-9.]+)$" % re.escape(key), f.read(), re.M)
    except OSError:
        return None
    return float(m.group(1)) if m else None


def measured_traffic_per_photon_step():
    return profile_figure("hbm_bytes_per_photon_step")


def algorithmic_bytes(counters, n_hist):
    return (SOURCE_BYTES * n_hist + REGION_BYTES * counters["region_gathers"]
            + TABLE_ROW_BYTES * (ROWS_PER_PHOTON_STEP * counters["photon_steps"]
                                 + ROWS_PER_ELECTRON_STEP * counters["electron_steps"])
            + DEPOSIT_BYTES * counters["deposits"])


def cpu_baseline(wl, budget_s, threads):
    """The oracle, built here for this host's CPU (gcc -O3 -march=native, no
    fast-math), on `threads` OpenMP threads over a bounded sample of the
    workload (about budget_s seconds)."""
    import subprocess
    import tempfile
    from oracle import oracle
    build = "gcc -O3 -march=native"
    try:
        oracle.use_library(oracle.build_native(tempfile.mkdtemp(prefix="gegs_oracle_")))
    except (OSError, subprocess.CalledProcessError):
        oracle.use_library(None)
        build = "gcc -O3 -march=x86-64-v3, prebuilt (no compiler on this host)"
    probe = wl.source(50000 * threads, 991)
    t0 = time.perf_counter()
    oracle.run(wl, probe, seed=5, threads=threads)
    rate0 = len(probe) / (time.perf_counter() - t0)
    n = int(max(20000, min(50_000_000, rate0 * budget_s)))
    sample = wl.source(n, 992)
    t0 = time.perf_counter()
    oracle.run(wl, sample, seed=6, threads=threads)
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="histories/s", cores=threads, kind="port",
                sample="%d histories of the same CT workload (seed 992), the oracle "
                       "(C restatement, %s, OpenMP over histories), %.1f s" % (n, build, dt))


def cpu_detail(wl, threads, seconds=4.0):
    """The oracle beside the headline baseline: one core on the same CT
    workload, and all cores on the 60^3 water phantom (1.25 MeV parallel
    beam), each on a bounded sample of about `seconds`."""
    from oracle import oracle
    out = {"nproc": os.cpu_count()}
    for name, w, th in (("ct_1core", wl, 1), ("water60_%dcores" % threads, None, threads)):
        if w is None:
            from gegs_amd import workloads
            w = workloads.water()
        probe = w.source(20000 * th, 993)
        t0 = time.perf_counter()
        oracle.run(w, probe, seed=7, threads=th)
        rate0 = len(probe) / (time.perf_counter() - t0)
        n = int(max(10000, min(20_000_000, rate0 * seconds)))
        sample = w.source(n, 994)
        t0 = time.perf_counter()
        oracle.run(w, sample, seed=8, threads=th)
        dt = time.perf_counter() - t0
        out[name] = dict(histories_per_s=n / dt, histories=n, seconds=dt, cores=th)
    return out


def atomic_peak(dose, deposits_per_s, reps=6):
    """FP64 atomic adds per second that the device sustains on the dose
    array when the target voxels follow the measured dose distribution (the
    observed conflict pattern), with the tally's own hardware atomic
    (gegs_atomic_probe), beside the event loop's deposits per second."""
    from gegs_amd import native
    p = np.clip(np.asarray(dose, dtype=np.float64), 0.0, None)
    p = p / p.sum()
    idx = np.random.default_rng(5).choice(len(p), size=1 << 24, p=p).astype(np.uint32)
    ctx = native.Context(0)
    try:
        rate = ctx.atomic_probe(idx, len(p), reps=reps)
    finally:
        ctx.close()
    return dict(peak_fp64_atomics_per_s=rate, event_loop_deposits_per_s=deposits_per_s,
                frac=deposits_per_s / rate,
                method="gegs_atomic_probe: %d FP64 atomic adds over the %d voxels, drawn from the dose "
                       "distribution of the run" % (reps * len(idx), len(p)))


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def add_counters(acc, c):
    if acc is None:
        return dict(c)
    return {k: (acc[k] + c[k] if not isinstance(c[k], list) else [a + b for a, b in zip(acc[k], c[k])])
            for k in acc}


def run_method(ctx, method, steps, warmup, b, seed, kernel_timing=False):
    """Times `steps` steps of b records each, as one call of the method with
    `steps` statistical batches (the plain kernel launches batch after
    batch; the event path carries all batches in one loop and tallies each
    in its own dose row).  With kernel_timing the event path launches its
    sweeps directly on the stream with a HIP event pair around every photon
    kernel launch (graph replays cannot carry timed events).  Returns
    (elapsed, dose, dose2, counters, timing)."""
    if warmup:
        # same batch count as the timed call, so that its graph is reused
        ctx.run(seed=seed, n_batches=steps, method=method, first=0, count=warmup * b)
    if kernel_timing:
        ctx.set_kernel_timing(True)
    _sync()
    t0 = time.perf_counter()
    r = ctx.run(seed=seed, n_batches=steps, method=method, first=warmup * b, count=steps * b)
    _sync()
    elapsed = time.perf_counter() - t0
    timing = None
    if kernel_timing:
        timing = ctx.kernel_timing()
        ctx.set_kernel_timing(False)
    return elapsed, r["dose"], r["dose2"], r["counters"], timing

def plan_sigma(dose, dose2, n_batches):
    """Mean relative uncertainty over voxels above 50% of Dmax of a dose
    summed over n_batches batches (dose2 = sum of squared batch doses)."""
    var = np.maximum(dose2 - dose * dose / n_batches, 0.0) * n_batches / (n_batches - 1)
    sel = dose > 0.5 * dose.max()
    return float(np.mean(np.sqrt(var[sel]) / dose[sel])), int(sel.sum()), var


def mlc_plan_bench(n_total, n_batches, seed, autotune_first=True, fixed_tuning=None):
    """The 5-field MLC plan, each method through its own interface: the
    plain kernel runs field after field (its leaf set and gantry loaded, one
    run of n_batches batches per field); the batched path runs the five
    fields in one event loop whose every batch holds 1/n_batches of every
    field.  Both transport the same histories with the same random streams,
    so the plan dose is the same; the uncertainty comes from each run's
    batch variance (summed over the fields for the plain runs)."""
    from gegs_amd import autotune, native, workloads
    plan = workloads.mlc_plan()
    src = plan.source(n_total, 3)
    ctx = native.Context(0)
    ctx.load_workload(plan.workload)
    ctx.set_source(src)
    tuning = fixed_tuning
    if fixed_tuning:
        ctx.set_tuning(fixed_tuning)
    elif autotune_first:
        ctx.set_leaf_positions(plan.leaf_positions[0])
        ctx.set_field(plan.fields[0])
        tuning, _, _ = autotune.autotune(ctx, 0, min(int(plan.counts(n_total)[0]), 1_000_000), seed=seed, repeat=3)
    jobs = plan.jobs(n_total)
    out = {}
    # warm-up: graph capture and first launches on a small slice
    ctx.set_leaf_positions(plan.leaf_positions)
    ctx.run_batched(plan.jobs(n_total, 0, 100), 1, seed=seed, n_batches=1)
    ctx.set_leaf_positions(plan.leaf_positions[0])
    ctx.set_field(plan.fields[0])
    ctx.run(seed=seed, n_batches=1, method="plain", first=0, count=min(100_000, n_total))
    for method in ("plain", "batched"):
        _sync()
        t0 = time.perf_counter()
        if method == "batched":
            ctx.set_leaf_positions(plan.leaf_positions)
            r = ctx.run_batched(jobs, 1, seed=seed, n_batches=n_batches)
            dose, var = r["dose"][0], None
            sig, nsel, _ = plan_sigma(dose, r["dose2"][0], n_batches)
            counters = r["counters"]
        else:
            dose, var, counters = None, None, None
            for jb in jobs:
                ctx.set_leaf_positions(plan.leaf_positions[jb["leaf_set"]])
                ctx.set_field(jb["field"])
                r = ctx.run(seed=seed, n_batches=n_batches, method="plain", first=jb["first"], count=jb["count"])
                _, _, v = plan_sigma(r["dose"], r["dose2"], n_batches)
                dose = r["dose"] if dose is None else dose + r["dose"]
                var = v if var is None else var + v
                counters = add_counters(counters, r["counters"])
        _sync()
        el = time.perf_counter() - t0
        if method == "plain":
            sel = dose > 0.5 * dose.max()
            sig, nsel = float(np.mean(np.sqrt(var[sel]) / dose[sel])), int(sel.sum())
        out[method] = dict(histories_per_s=n_total / el, seconds=el, mean_rel_sigma=sig, voxels=nsel,
                           dose_efficiency=1.0 / (sig * sig * el), photon_steps=counters["photon_steps"],
                           electron_steps=counters["electron_steps"], dose_sum=float(dose.sum()))
    ctx.close()
    out["batched_vs_plain_efficiency"] = out["batched"]["dose_efficiency"] / out["plain"]["dose_efficiency"]
    out["batched_vs_plain_time"] = out["plain"]["seconds"] / out["batched"]["seconds"]
    out.update(workload="ct_128x128x100_mlc_5field_gantry_0_72_144_216_288", histories=n_total,
               n_batches=n_batches, field_weights=[float(x) for x in plan.weights], tuning=tuning,
               source="egsphsp MODE0 list decoded at z = 16.44 cm, %.1f%% electrons"
                      % (100.0 * float(np.mean(src["charge"] != 0))))
    return out


def mlc_plan_at_one_percent(n_target, n_records, n_batches, seed, tuning, plain_efficiency):
    """The batched 5-field MLC plan run to the stated uncertainty: its
    n_records phase-space records recycled R times (R even, each pass with
    its own Philox history range, as the reference's nRepeat recycling), in
    two independent halves of R/2 passes and n_batches/2 statistical
    batches each.  sigma comes from the pooled n_batches batches; its spread
    from the two halves' own estimates, each scaled to the whole run.  The
    plain kernel's dose efficiency 1/(sigma^2 T) does not depend on the
    number of histories (sigma^2 ~ 1/N, T ~ N), so the ratio takes the plain
    runs' efficiency at their own scale."""
    import math
    from gegs_amd import native, workloads
    plan = workloads.mlc_plan()
    src = plan.source(n_records, 3)
    reps = 2 * int(math.ceil(n_target / (2.0 * n_records)))
    ctx = native.Context(0)
    ctx.load_workload(plan.workload)
    ctx.set_source(src)
    if tuning:
        ctx.set_tuning(tuning)
    ctx.set_leaf_positions(plan.leaf_positions)
    base = plan.jobs(n_records)
    half_b = n_batches // 2
    # warm-up: graph capture on a small slice
    ctx.run_batched(plan.jobs(n_records, 0, 100), 1, seed=seed, n_batches=half_b)
    halves = []
    for h in range(2):
        jobs = [dict(jb, history_offset=r * n_records + jb["first"])
                for r in range(h * reps // 2, (h + 1) * reps // 2) for jb in base]
        _sync()
        t0 = time.perf_counter()
        out = ctx.run_batched(jobs, 1, seed=seed, n_batches=half_b)
        _sync()
        halves.append((time.perf_counter() - t0, out["dose"][0], out["dose2"][0], out["counters"]["overflow"]))
    ctx.close()
    t = halves[0][0] + halves[1][0]
    dose = halves[0][1] + halves[1][1]
    dose2 = halves[0][2] + halves[1][2]
    sig, nsel, _ = plan_sigma(dose, dose2, 2 * half_b)
    half_sig = [plan_sigma(hv[1], hv[2], half_b)[0] / math.sqrt(2.0) for hv in halves]
    n_hist = reps * n_records
    eff = 1.0 / (sig * sig * t)
    return dict(histories=n_hist, record_passes=reps, n_batches=2 * half_b, seconds=t, histories_per_s=n_hist / t,
                mean_rel_sigma=sig, voxels=nsel, dose_efficiency=eff,
                sigma_from_each_half_scaled=half_sig,
                sigma_spread=abs(half_sig[0] - half_sig[1]) / sig,
                overflow=int(halves[0][3] + halves[1][3]),
                batched_vs_plain_efficiency=eff / plain_efficiency if plain_efficiency else None,
                plain_reference="plain per-field runs of mlc_plan at their own scale (efficiency is N-independent)")


def adjoint_rate(n_hist, seed, repeat=2):
    """Gradients/s of the beamlet adjoint on the dose-optimisation workload:
    the 128^3 CT at 2.5 mm, 5 fields (gantry 0/72/144/216/288) x 20x20
    beamlets of 5 mm, w ~ U(0.5, 1.5), L = sum m (D - P)^2 over a 3 cm target
    sphere and a 4-5 cm OAR shell; one gradient = one replay of n_hist
    histories."""
    from gegs_amd import beamlets, native, workloads
    wl = workloads.ct_cube()
    fields = [workloads.make_field(True, wl.field.source_to_isocenter, g) for g in workloads.C3_GANTRY]
    per = n_hist // len(fields)
    src = wl.source(per * len(fields), 777)
    ctx = native.Context(0)
    ctx.load_workload(wl)
    ctx.set_source(src)
    plan = beamlets.BeamletPlan(ctx, fields, per, n_side=20, pitch=0.5, seed=seed, n_batches=2, src=src)
    w = np.random.default_rng(4).uniform(0.5, 1.5, plan.n_beamlets)
    d = plan.forward(w)
    m, p = workloads.objective(wl.phantom, d)
    dLdD = 2.0 * m * (d - p)
    plan.adjoint(dLdD)                      # graph capture / warm caches
    best = None
    for _ in range(repeat):
        _sync()
        t0 = time.perf_counter()
        g = plan.adjoint(dLdD)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    ctx.close()
    return dict(gradients_per_s=1.0 / best, ms_per_gradient=best * 1e3, histories_per_gradient=per * len(fields),
                beamlets=plan.n_beamlets, fields=len(fields), nonzero=int(np.count_nonzero(g)),
                workload="%s, target sphere r 3 cm + OAR shell 4-5 cm" % wl.name)


def plan_queue_bench(total, n_batches, seed, tuning=None):
    """16 plans of 1-9 fields (budgets in [1e6, 2e7] summing to `total`) on
    the 256 x 256 x 180 CT at 2 mm: all 93 fields in one batched run (one
    dose buffer per plan) against the plain kernel field after field.  Same
    histories and streams on both sides; the per-plan uncertainty comes from
    each method's batch variance.  Then the queue's beamlet-weight gradient
    (20 x 20 beamlets of 5 mm per field): one replay of all `total`
    histories with one dL/dD row per plan (target sphere and OAR shell of
    each plan's own dose), dL/dD and the gradient in device memory."""
    import torch
    from gegs_amd import beamlets, native, workloads
    mp = workloads.multi_plan(total=total)
    src = mp.source(5)
    jobs = mp.jobs()
    nplan = len(mp.plans)
    ctx = native.Context(0)
    ctx.load_workload(mp.workload)
    ctx.set_source(src)
    queue = beamlets.BeamletPlan.from_jobs(ctx, jobs, nplan, src, n_side=20, pitch=0.5, seed=seed, n_batches=2)
    del src
    if tuning:
        ctx.set_tuning(tuning)
    # warm-up: graph capture and first launches on small slices
    small = [dict(jb, count=min(jb["count"], 2000)) for jb in jobs]
    ctx.run_batched(small, nplan, seed=seed, n_batches=1)
    ctx.set_field(jobs[0]["field"])
    ctx.run(seed=seed, n_batches=1, method="plain", first=0, count=100_000)
    _sync()
    t0 = time.perf_counter()
    b = ctx.run_batched(jobs, nplan, seed=seed, n_batches=n_batches)
    _sync()
    t_b = time.perf_counter() - t0
    dose = np.zeros_like(b["dose"])
    var = np.zeros_like(b["dose"])
    _sync()
    t0 = time.perf_counter()
    for jb in jobs:
        ctx.set_field(jb["field"])
        r = ctx.run(seed=seed, n_batches=n_batches, method="plain", first=jb["first"], count=jb["count"])
        _, _, v = plan_sigma(r["dose"], r["dose2"], n_batches)
        dose[jb["plan"]] += r["dose"]
        var[jb["plan"]] += v
    _sync()
    t_p = time.perf_counter() - t0
    target, oar = workloads.target_and_oar(mp.workload.phantom)
    m = (target | oar).astype(np.float64)
    P = np.where(target[None, :], 1.2 * b["dose"][:, target].mean(axis=1)[:, None], 0.0)
    dLdD = torch.from_numpy(2.0 * m[None, :] * (b["dose"] - P)).cuda()
    del P
    queue.adjoint_device(dLdD)              # graph capture / warm caches
    _sync()
    t0 = time.perf_counter()
    g = queue.adjoint_device(dLdD)
    _sync()
    t_a = time.perf_counter() - t0
    nonzero = int(torch.count_nonzero(g).item())
    del dLdD, g
    ctx.close()
    ratio, sig_b, sig_p = [], [], []
    for k in range(nplan):
        sb, _, _ = plan_sigma(b["dose"][k], b["dose2"][k], n_batches)
        sel = dose[k] > 0.5 * dose[k].max()
        sp = float(np.mean(np.sqrt(var[k][sel]) / dose[k][sel]))
        sig_b.append(sb)
        sig_p.append(sp)
        ratio.append((sp * sp * t_p) / (sb * sb * t_b))
    rel = float(np.max(np.abs(dose - b["dose"])) / np.max(np.abs(dose)))
    return dict(workload=mp.name, plans=nplan, fields=len(jobs), histories=mp.n_histories, n_batches=n_batches,
                batched=dict(seconds=t_b, histories_per_s=mp.n_histories / t_b),
                plain=dict(seconds=t_p, histories_per_s=mp.n_histories / t_p),
                batched_vs_plain_time=t_p / t_b, batched_vs_plain_efficiency_mean_over_plans=float(np.mean(ratio)),
                mean_rel_sigma_batched=float(np.mean(sig_b)), mean_rel_sigma_plain=float(np.mean(sig_p)),
                max_rel_dose_difference=rel,
                adjoint=dict(seconds=t_a, queue_gradients_per_s=1.0 / t_a, plan_gradients_per_s=nplan / t_a,
                             histories_per_s=mp.n_histories / t_a, beamlets=queue.n_beamlets, nonzero=nonzero,
                             objective="per plan: target sphere r 3 cm (1.2 x its mean dose) + OAR shell 4-5 cm"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories-per-step", type=int, default=1_000_000)
    ap.add_argument("--methods", default="plain,event", help="comma list of plain,event")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default=None,
                    help="JSON dict of event-path tuning fields to use instead of autotuning")
    ap.add_argument("--adjoint-histories", type=int, default=10_000_000,
                    help="histories per adjoint gradient (0: skip)")
    ap.add_argument("--mlc-plan-histories", type=int, default=20_000_000,
                    help="histories of the 5-field MLC plan comparison (0: skip)")
    ap.add_argument("--mlc-plan-1pct-histories", type=int, default=1_700_000_000,
                    help="histories of the batched 5-field MLC plan run to about 1%% mean uncertainty (0: skip)")
    ap.add_argument("--plan-queue-histories", type=int, default=100_000_000,
                    help="histories of the 16-plan queue comparison on the 256x256x180 CT (0: skip)")
    ap.add_argument("--seed", type=int, default=20240)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.copy_bytes:
        import torch
        a = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(histories shard across GPUs with one FP64 dose reduction; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from gegs_amd import autotune, native, workloads
    wl = workloads.ct()
    b = args.histories_per_step
    n_total = (args.steps + args.warmup) * b
    src = wl.source(n_total, 4242)
    ctx = native.Context(0)
    ctx.load_workload(wl)
    ctx.set_source(src)          # resident in HBM before any timed step

    methods = [m.strip() for m in args.methods.split(",") if m.strip()]
    tuning = None
    if args.tuning:
        tuning = json.loads(args.tuning)
        ctx.set_tuning(tuning)
    elif "event" in methods and not args.no_autotune:
        # tuned on the warmup records, before the timed region
        tuning, _, _ = autotune.autotune(ctx, 0, min(max(args.warmup, 1) * b, 2_000_000), seed=args.seed, repeat=3)
    results = {}
    for method in methods:
        elapsed, dose, dose2, cnt, _ = run_method(ctx, method, args.steps, args.warmup, b, args.seed)
        hist = args.steps * b
        sig, nsel, _ = plan_sigma(dose, dose2, args.steps)
        eff = 1.0 / (sig * sig * elapsed)
        res = dict(value=hist / elapsed, ms_per_step=elapsed / args.steps * 1e3, sigma=sig, voxels=nsel,
                   efficiency=eff, counters=cnt, bytes=algorithmic_bytes(cnt, hist), dose=dose,
                   deposits_per_s=cnt["deposits"] / elapsed)
        if method == "event":
            # the same steps again with every photon kernel launch timed by
            # HIP events on the library's stream
            el2, _, _, cnt2, timing = run_method(ctx, method, args.steps, 0, b, args.seed, kernel_timing=True)
            if timing and timing[1]:
                photon_ms, launches = timing
                pbytes = PHOTON_STEP_BYTES * cnt2["photon_steps"]
                per_step = measured_traffic_per_photon_step()
                res["roofline"] = dict(kernel="photon_kernel", launches=launches,
                                       traffic=(per_step * cnt2["photon_steps"] / launches) if per_step else None,
                                       avg_launch_us=photon_ms / launches * 1e3,
                                       algorithmic_bytes_per_launch=pbytes / launches,
                                       achieved=pbytes / (photon_ms * 1e-3) / 1e9,
                                       photon_kernel_share=photon_ms * 1e-3 / el2,
                                       timed_pass_histories_per_s=hist / el2)
                # the photon kernel alone: the same steps with the electron
                # side of each sweep after the photon side instead of beside it
                tun = ctx.get_tuning()
                if tun.get("sweep_streams", 2) == 2:
                    ctx.set_tuning(dict(tun, sweep_streams=1))
                    el3, _, _, cnt3, timing3 = run_method(ctx, method, args.steps, 0, b, args.seed,
                                                          kernel_timing=True)
                    ctx.set_tuning(tun)
                    if timing3 and timing3[1]:
                        ms3, n3 = timing3
                        b3 = PHOTON_STEP_BYTES * cnt3["photon_steps"]
                        res["roofline"]["isolated"] = dict(
                            sweep_streams=1, launches=n3, avg_launch_us=ms3 / n3 * 1e3,
                            achieved=b3 / (ms3 * 1e-3) / 1e9, frac=b3 / (ms3 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                            timed_pass_histories_per_s=hist / el3)
        results[method] = res

    head = "event" if "event" in results else methods[0]
    r = results[head]
    cpu = None
    cpu_more = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(wl, args.cpu_seconds, threads)
        cpu_more = cpu_detail(wl, threads)
    if tuning is not None:
        # the event loop's whole effective tuning, defaults included
        eff = ctx.get_tuning()
        tuning = dict(tuning, sweep_streams=eff["sweep_streams"], tail_particles=eff["tail_particles"])
    ctx.close()
    atomics = atomic_peak(r["dose"], r["deposits_per_s"]) if head == "event" else None
    adj = None
    if args.adjoint_histories:
        adj = adjoint_rate(args.adjoint_histories, args.seed)
    mlc_plan = None
    if args.mlc_plan_histories:
        mlc_plan = mlc_plan_bench(args.mlc_plan_histories, args.steps, args.seed, not args.no_autotune,
                                  json.loads(args.tuning) if args.tuning else None)
    one_pct = None
    if args.mlc_plan_1pct_histories and mlc_plan:
        one_pct = mlc_plan_at_one_percent(args.mlc_plan_1pct_histories, args.mlc_plan_histories, 100, args.seed,
                                          mlc_plan.get("tuning"), mlc_plan["plain"]["dose_efficiency"])
        mlc_plan["at_1pct"] = one_pct
    queue = None
    if args.plan_queue_histories:
        queue = plan_queue_bench(args.plan_queue_histories, args.steps, args.seed, tuning)
    if "roofline" in r:
        rf = r["roofline"]
        roof = {"bound": "hbm", "achieved": rf["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": rf["achieved"] / HBM_PEAK_GBS, "traffic": rf["traffic"], "kernel": rf["kernel"],
                "traffic_source": "profiles/photon_kernel_ct128_event.md: rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 "
                                  "(1 GiB copy calibration) per photon step, times this run's photon steps per launch",
                "avg_launch_us": rf["avg_launch_us"], "launches": rf["launches"],
                "algorithmic_bytes_per_launch": rf["algorithmic_bytes_per_launch"],
                "algorithmic_bytes_per_photon_step": PHOTON_STEP_BYTES,
                "photon_kernel_share": rf["photon_kernel_share"],
                "timed_pass_histories_per_s": rf["timed_pass_histories_per_s"],
                "timing": "HIP events around every photon kernel launch over a second pass of the same steps "
                          "(sweeps launched on the stream; the value pass replays graphs); the photon kernel "
                          "shares the GPU with the electron side of its sweep, which runs on a second stream",
                "isolated": rf.get("isolated"),
                "valu_issue_frac": profile_figure("valu_issue_frac"),
                "limiter": "VALU issue: the SIMDs issue VALU most of the kernel's cycles (valu_issue_frac, "
                           "SQ_ACTIVE_INST_VALU in the committed profile); real HBM traffic is traffic/launch"}
    else:
        roof = None
    line = {
        "metric": "histories/s + dose eff. 1/(sigma^2 T) at 1% unc., batched vs plain; adjoint grad/s",
        "value": r["value"],
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "ct_128x128x100_2.5mm_6MV_point_source_SSD100_10x10",
            "histories_per_step": b,
            "histories": args.steps * b,
            "method": "event-queued, batched" if head == "event" else head,
            "ecut_mev": 0.7, "pcut_mev": 0.01,
            "tallies": "fp64",
        },
        "dose_efficiency": r["efficiency"],
        "dose_efficiency_unit": "1/(sigma^2 s), sigma = mean relative uncertainty over D > 50% Dmax",
        "mean_rel_sigma": r["sigma"],
        "histories_for_1pct": args.steps * b * (r["sigma"] / 0.01) ** 2,
        "methods": {m: dict(histories_per_s=v["value"], dose_efficiency=v["efficiency"],
                            ms_per_step=v["ms_per_step"],
                            algorithmic_bytes_per_history=v["bytes"] / (args.steps * b),
                            photon_steps_per_history=v["counters"]["photon_steps"] / (args.steps * b),
                            electron_steps_per_history=v["counters"]["electron_steps"] / (args.steps * b))
                    for m, v in results.items()},
        "tuning": tuning,
        "roofline": roof,
        "adjoint": adj,
        "mlc_plan": mlc_plan,
        "plan_queue": queue,
        "cpu_baseline": cpu,
        "cpu_baseline_detail": cpu_more,
        "fp64_atomics": atomics,
    }
    if mlc_plan:
        # the batching target is stated on the 5-field MLC plan
        line["batched_vs_plain_efficiency"] = mlc_plan["batched_vs_plain_efficiency"]
        line["batched_vs_plain_workload"] = mlc_plan["workload"]
    if "plain" in results and "event" in results:
        line["event_vs_plain_efficiency_single_field_ct"] = results["event"]["efficiency"] / results["plain"]["efficiency"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
