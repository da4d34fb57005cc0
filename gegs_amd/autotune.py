"""Autotuner of the event-queued method's launch shape and queue policy.

The reference fixes its stack sizes, block and grid shapes at compile time
(CUDA_EGS.h, init.cu:31-57).  Here the throughput of the event loop depends
on the workload: how long photon flights are (voxel count along the beam,
density), how many electrons each history sets in motion, how many
histories a batch holds.  ``autotune`` measures histories/s of short runs on
a sample of the workload's own source records and walks the tuning space one
coordinate at a time (photon chunk, electron chunk, refill threshold and
reservation, flight block size, sweeps per graph), keeping a change only when
it is faster by more than the run-to-run noise.  The injection size is not
searched: its effect shows only in runs longer than one injection, which a
short trial is not; it stays at 6 Mi primaries (the fastest of 1-8 Mi on
1e7-history runs of the CT workload).  Every trial transports real histories with the
production kernels; the tuning never changes results, only speed (the event
path reproduces the plain kernel history by history for every tuning).
"""
from __future__ import annotations

import time

SPACE = {
    "photon_chunk": [16, 32, 64, 128],
    "electron_chunk": [8, 16, 32, 64],
    "refill_min": [8, 16, 32, 48],
    "refill_slots": [64, 256, 1024],
    "flight_block": [256, 512],
    "sweeps_per_graph": [2, 4, 8],
}

START = dict(photon_chunk=64, electron_chunk=16, refill_min=16, refill_slots=64, flight_block=256,
             inject=6 << 20, threshold=6 << 20, sweeps_per_graph=4)


def _rate(ctx, tuning, first, count, seed, repeat):
    ctx.set_tuning(tuning)
    best = 0.0
    for _ in range(repeat):
        t0 = time.perf_counter()
        ctx.run(seed=seed, n_batches=1, method="event", first=first, count=count)
        best = max(best, count / (time.perf_counter() - t0))
    return best


def autotune(ctx, first, count, seed=1, repeat=2, tol=0.02, rounds=1, log=None):
    """Tune the event method on source records [first, first+count) of the
    loaded context.  Returns (tuning dict, histories/s of the best trial,
    list of (tuning, histories/s) trials).  The context keeps the best tuning."""
    cur = dict(START)
    trials = []
    _rate(ctx, cur, first, count, seed, 1)            # graph capture, warm caches
    best = _rate(ctx, cur, first, count, seed, repeat)
    trials.append((dict(cur), best))
    for _ in range(rounds):
        for key, values in SPACE.items():
            for v in values:
                if v == cur[key]:
                    continue
                t = dict(cur)
                t[key] = v
                if key == "inject":
                    t["threshold"] = v
                r = _rate(ctx, t, first, count, seed, repeat)
                trials.append((dict(t), r))
                if log:
                    log("autotune %s=%s: %.3g histories/s" % (key, v, r))
                if r > best * (1.0 + tol):
                    best, cur = r, t
    ctx.set_tuning(cur)
    return cur, best, trials
