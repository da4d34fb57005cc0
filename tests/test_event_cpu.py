"""The event-queued path's own kernels, compiled for the CPU, against the
oracle: per-history energy and step counts must be identical.

tests/cpu_event/event_cpu.cpp includes gegs_amd/csrc/event.hpp unchanged and
compiles it as host C++ with AddressSanitizer (host code only) against
tests/cpu_event/include/hip/hip_runtime.h, where every thread runs as a
wavefront with one active lane.  This checks, without a GPU, the bank
layout and packing, injection, suspension and resumption of flights
(chunked photon walks, electron step state), queue routing between the ten
event banks, the module chain jaws -> MLC -> phantom, the Philox stream
bookkeeping and continuous batching (two statistical batches in one pass,
each deposit in its history's batch row); the LDS array is sized exactly to the staged bytes so that an
out-of-range read of the staged tables fails the run.  The second build
(tests/cpu_event/include_simt) runs the same kernels with 64-lane
wavefronts, for what happens between the lanes of one wavefront."""
import ctypes
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from gegs_amd import workloads
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CLANG = next((c for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++")) if c and os.path.exists(c)), None)

pytestmark = pytest.mark.skipif(CLANG is None, reason="no clang++ for the CPU build of the event path")


def lds_floats(wl):
    """Bytes the flight kernels stage (env_lds_bytes with the record table
    in LDS, one job) divided by 4."""
    reg = np.ascontiguousarray(wl.regions)
    nclass = len(np.unique(reg.view(np.dtype((np.void, reg.dtype.itemsize)))))
    ph = wl.phantom
    nmed = wl.tables.c.nmed
    assert nclass <= 2048
    b = 32 + (ph.nx + ph.ny + ph.nz + 6) * 4 + nclass * 16 + nmed * 8 + 2 * nmed * 200 * 8
    return b // 4


def build(tmp, floats, simt=False):
    exe = os.path.join(tmp, "event_cpu_%s%d" % ("simt_" if simt else "", floats))
    inc = "include_simt" if simt else "include"
    if not os.path.exists(exe):
        subprocess.check_call([CLANG, "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
                               "-ffp-contract=off", "-I", os.path.join(HERE, "cpu_event", inc),
                               "-DLDS_FLOATS=%d" % floats] + (["-DGEGS_SIMT"] if simt else []) +
                              ["-o", exe, os.path.join(HERE, "cpu_event", "event_cpu.cpp")])
    return exe


def dump(wl, src, seed, d, nbatch, pchunk, echunk, cap=1 << 16, flight_block=256, refill_min=16, refill_slots=64,
         flight_grid=4, inject=0, threshold=0, tail=0):
    os.makedirs(d, exist_ok=True)
    ph = wl.phantom
    first = wl.field.module_first
    mlc = getattr(wl, "mlc", None)
    have_mlc = 1 if (first == 3 or (first == 1 and mlc is not None)) else 0
    have_jaws = 1 if first == 1 else 0

    def put(name, data):
        with open(os.path.join(d, name), "wb") as f:
            f.write(data)

    put("meta.bin", struct.pack("<10I2Q7I", ph.nx, ph.ny, ph.nz, len(src), nbatch, pchunk, echunk, have_mlc,
                                have_jaws, cap, seed, 0, flight_block, refill_min, refill_slots, flight_grid,
                                inject, threshold, tail))
    put("tables.bin", ctypes.string_at(ctypes.addressof(wl.tables.c), ctypes.sizeof(wl.tables.c)))
    put("regions.bin", np.ascontiguousarray(wl.regions).tobytes())
    for n, b in (("xb", ph.xb), ("yb", ph.yb), ("zb", ph.zb)):
        put(n + ".bin", np.ascontiguousarray(b, dtype=np.float32).tobytes())
    put("field.bin", ctypes.string_at(ctypes.addressof(wl.field), ctypes.sizeof(wl.field)))
    put("src.bin", np.ascontiguousarray(src).tobytes())
    sid = wl.field.source_to_isocenter
    if have_mlc:
        put("mlc.bin", ctypes.string_at(ctypes.addressof(mlc), ctypes.sizeof(mlc)))
        put("endnodes.bin", oracle.leaf_ends(mlc, wl.leaf_positions, sid).tobytes())
    if have_jaws:
        put("jaws.bin", oracle.jaw_polys(wl.jaws, wl.jaw_positions, sid).tobytes())


CASES = {
    "water": (lambda: workloads.water(), 1500, 32, 16),
    "ct": (lambda: workloads.ct(), 400, 32, 16),
    "ct_suspend_every_step": (lambda: workloads.ct(), 150, 1, 1),
    "ct_mlc": (lambda: workloads.ct_mlc(), 2500, 32, 16),
    "ct_jaws_mlc": (lambda: workloads.ct_jaws_mlc(), 2500, 8, 4),
    # the tail launch takes over once at most 300 particles are in flight
    "ct_tail": (lambda: workloads.ct(), 400, 32, 16, 300),
    "ct_mlc_tail": (lambda: workloads.ct_mlc(), 2500, 16, 8, 300),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_event_path_cpu_build_matches_oracle(case, tmp_path):
    make, n, pchunk, echunk = CASES[case][:4]
    tail = CASES[case][4] if len(CASES[case]) > 4 else 0
    wl = make()
    src = wl.source(n, 11)
    d = str(tmp_path / case)
    dump(wl, src, 77, d, 2, pchunk, echunk, tail=tail)
    exe = build(str(tmp_path), lds_floats(wl))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
    r = subprocess.run([exe, d], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    assert ("tail" in r.stdout) == (tail > 0), r.stdout
    edep = np.fromfile(os.path.join(d, "out_edep.bin"), dtype=np.float64)
    steps = np.fromfile(os.path.join(d, "out_steps.bin"), dtype=np.uint32)
    rows = np.fromfile(os.path.join(d, "out_dose.bin"), dtype=np.float64).reshape(2, -1)
    o = oracle.run(wl, src, seed=77, per_history=True, threads=4)
    assert np.array_equal(steps, o["steps"])
    assert np.allclose(edep, o["edep"], rtol=1e-12, atol=1e-12)
    assert np.any(edep > 0)
    # both batches ran in one pass; each deposit is in its history's batch row
    half = n // 2
    for b, (lo, hi) in enumerate(((0, half), (half, n))):
        ob = oracle.run(wl, src[lo:hi], seed=77, hist_offset=lo, threads=4)
        assert np.allclose(rows[b], ob["dose"], rtol=1e-11, atol=1e-12), b


# The same kernels with 64-lane wavefronts (tests/cpu_event/include_simt):
# grouped bank appends with mixed destinations, refill reservations shared by
# the idle lanes of a wavefront, held photon pushes and the look-ahead vote,
# with the GPU's launch shape (256-thread flight blocks, refill after 16 idle
# lanes, 64-slot reservations).  The stand-in stops the run on a divergent
# whole-wavefront operation or a read of an inactive lane; AddressSanitizer
# on any access outside the banks, the counters or the staged LDS.
# (photon chunk, electron chunk, refill_min, refill_slots, flight blocks, inject, threshold)
SIMT_CASES = {
    "ct_seed7_chunk64": (lambda: workloads.ct(), 1500, 7, (64, 16, 16, 64, 6, 1024, 1024)),
    "ct_refill1_chunk8": (lambda: workloads.ct(), 1200, 7, (8, 4, 1, 64, 3, 512, 256)),
    "ct_mlc": (lambda: workloads.ct_mlc(), 2000, 7, (64, 16, 16, 128, 4, 1024, 2048)),
    "ct_jaws_mlc_chunk4": (lambda: workloads.ct_jaws_mlc(), 1500, 7, (4, 2, 16, 64, 4, 768, 512)),
    "ct_tail_chunk16": (lambda: workloads.ct(), 1200, 7, (16, 8, 16, 64, 4, 512, 512), 700),
}


@pytest.mark.parametrize("case", sorted(SIMT_CASES))
def test_event_path_64_lane_build_matches_oracle(case, tmp_path):
    make, n, seed, (pchunk, echunk, rmin, rslots, fgrid, inject, thr) = SIMT_CASES[case][:4]
    tail = SIMT_CASES[case][4] if len(SIMT_CASES[case]) > 4 else 0
    wl = make()
    src = wl.source(n, 11)
    d = str(tmp_path / case)
    dump(wl, src, seed, d, 2, pchunk, echunk, flight_block=256, refill_min=rmin, refill_slots=rslots,
         flight_grid=fgrid, inject=inject, threshold=thr, tail=tail)
    exe = build(str(tmp_path), lds_floats(wl), simt=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:detect_stack_use_after_return=0")
    r = subprocess.run([exe, d], capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-3000:])
    assert ("tail" in r.stdout) == (tail > 0), r.stdout
    edep = np.fromfile(os.path.join(d, "out_edep.bin"), dtype=np.float64)
    steps = np.fromfile(os.path.join(d, "out_steps.bin"), dtype=np.uint32)
    rows = np.fromfile(os.path.join(d, "out_dose.bin"), dtype=np.float64).reshape(2, -1)
    o = oracle.run(wl, src, seed=seed, per_history=True, threads=4)
    assert np.array_equal(steps, o["steps"])
    assert np.allclose(edep, o["edep"], rtol=1e-12, atol=1e-12)
    assert np.allclose(rows.sum(axis=0), o["dose"], rtol=1e-11, atol=1e-12)


def _selftest_exe(tmp):
    exe = os.path.join(tmp, "simt_selftest")
    subprocess.check_call([CLANG, "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
                           "-ffp-contract=off", "-I", os.path.join(HERE, "cpu_event", "include_simt"), "-o", exe,
                           os.path.join(HERE, "cpu_event", "simt_selftest.cpp")])
    return exe


def test_64_lane_stand_in_and_wavefront_pieces(tmp_path):
    """The stand-in's shuffles and subset ballots, BankSink::push with mixed
    destinations from divergent lanes (every particle in its bank once) and
    Refill::take (every slot handed out once), under 64-lane wavefronts; and
    the stand-in stops a run on divergent whole-wavefront operations and on
    a readlane of an inactive lane."""
    exe = _selftest_exe(str(tmp_path))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0")
    r = subprocess.run([exe, "ok"], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert "0 failures" in r.stdout
    for mode, msg in (("divergent_uniform", "different wavefront operations"),
                      ("readlane_inactive", "outside the executing group")):
        r = subprocess.run([exe, mode], capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode != 0 and msg in r.stderr, (mode, r.stderr[-2000:])
