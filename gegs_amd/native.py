"""ctypes binding of the C ABI (include/gegs.h) in gegs_amd/libgegs_hip.so.

The product path has no CPU fallback: if the HIP library is missing or no
GPU is visible, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import hatch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgegs_hip.so")

# exported symbols declared in include/gegs.h
SYMBOLS = [
    "gegs_create", "gegs_destroy", "gegs_last_error", "gegs_sizeof_media_tables",
    "gegs_load_media", "gegs_load_phantom", "gegs_set_field", "gegs_set_source",
    "gegs_run_plain", "gegs_run_event", "gegs_set_history_tally", "gegs_history_tally",
    "gegs_get_counters", "gegs_last_kernel_ms", "gegs_sample_events", "gegs_atomic_probe",
    "gegs_ct_subsize", "gegs_ct_resample_grid", "gegs_ct_resample",
    "gegs_set_tuning", "gegs_get_tuning", "gegs_run_batched", "gegs_set_beamlets", "gegs_adjoint",
    "gegs_set_kernel_timing", "gegs_kernel_timing", "gegs_load_mlc", "gegs_set_leaf_positions",
    "gegs_load_jaws", "gegs_set_jaw_positions", "gegs_load_block_wedge", "gegs_set_fixes",
    "gegs_run_plain_noalloc", "gegs_run_batched_device", "gegs_adjoint_device",
    "gegs_set_beamlet_weights_device",
]

PARTICLE_DTYPE = np.dtype([("e", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("u", "<f4"),
                           ("v", "<f4"), ("w", "<f4"), ("wt", "<f4"), ("dnear", "<f4"),
                           ("region", "<u4"), ("charge", "<i4"), ("status", "<i4"),
                           ("lin", "<u8"), ("blk", "<u4"), ("pad", "<u4")])


class RunOptions(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("history_offset", ctypes.c_uint64),
                ("n_batches", ctypes.c_uint32), ("block_size", ctypes.c_uint32),
                ("grid_size", ctypes.c_uint32), ("stack_depth", ctypes.c_uint32),
                ("first", ctypes.c_uint64), ("count", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    """gegs_tuning: launch shape and queue policy of the event-queued method."""
    _fields_ = [("block", ctypes.c_uint32), ("grid", ctypes.c_uint32),
                ("photon_grid", ctypes.c_uint32), ("electron_grid", ctypes.c_uint32),
                ("photon_chunk", ctypes.c_uint32), ("electron_chunk", ctypes.c_uint32),
                ("inject", ctypes.c_uint32), ("threshold", ctypes.c_uint32),
                ("sweeps_per_graph", ctypes.c_uint32), ("flight_block", ctypes.c_uint32),
                ("refill_min", ctypes.c_uint32), ("refill_slots", ctypes.c_uint32),
                ("sweep_streams", ctypes.c_uint32), ("tail_particles", ctypes.c_uint32),
                ("capacity", ctypes.c_uint64)]

    FIELDS = ("block", "grid", "photon_grid", "electron_grid", "photon_chunk", "electron_chunk",
              "inject", "threshold", "sweeps_per_graph", "flight_block", "refill_min", "refill_slots",
              "sweep_streams", "tail_particles", "capacity")

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in self.FIELDS}


class Job(ctypes.Structure):
    """gegs_job: one field of a plan for gegs_run_batched."""
    _fields_ = [("field", ctypes.c_byte * 32), ("plan", ctypes.c_uint32),
                ("leaf_set", ctypes.c_uint32), ("first", ctypes.c_uint64), ("count", ctypes.c_uint64),
                ("history_offset", ctypes.c_uint64)]


class Counters(ctypes.Structure):
    _fields_ = [("photon_steps", ctypes.c_uint64), ("electron_steps", ctypes.c_uint64),
                ("region_gathers", ctypes.c_uint64), ("deposits", ctypes.c_uint64),
                ("interactions", ctypes.c_uint64 * 10), ("histories", ctypes.c_uint64),
                ("overflow", ctypes.c_uint64)]

    def as_dict(self):
        return dict(photon_steps=self.photon_steps, electron_steps=self.electron_steps,
                    region_gathers=self.region_gathers, deposits=self.deposits,
                    interactions=list(self.interactions), histories=self.histories,
                    overflow=self.overflow)


_lib = None


def load():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("gegs_amd/libgegs_hip.so is missing: run `python -m gegs_amd.build` "
                           "(the transport has no CPU fallback)")
    # PyTorch ships its own HIP runtime under the same soname
    # (libamdhip64.so.7).  Loading torch first makes the library bind to that
    # runtime, so that the process has one HIP runtime and torch tensors and
    # this library share the device (two runtimes do not both see the GPU).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.gegs_create.restype = vp
    L.gegs_create.argtypes = [ctypes.c_int]
    L.gegs_destroy.argtypes = [vp]
    L.gegs_last_error.restype = ctypes.c_char_p
    L.gegs_last_error.argtypes = [vp]
    L.gegs_sizeof_media_tables.restype = ctypes.c_size_t
    L.gegs_load_media.argtypes = [vp, vp]
    L.gegs_load_phantom.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp, vp]
    L.gegs_set_field.argtypes = [vp, vp]
    L.gegs_set_source.argtypes = [vp, vp, ctypes.c_uint64]
    L.gegs_run_plain.argtypes = [vp, vp, vp, vp]
    L.gegs_run_event.argtypes = [vp, vp, vp, vp]
    L.gegs_run_plain_noalloc.argtypes = [vp, vp, vp, vp]
    L.gegs_set_history_tally.argtypes = [vp, ctypes.c_int]
    L.gegs_history_tally.argtypes = [vp, vp, vp]
    L.gegs_get_counters.argtypes = [vp, vp]
    L.gegs_last_kernel_ms.argtypes = [vp, vp]
    L.gegs_sample_events.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint64, vp, ctypes.c_uint32, vp, vp]
    L.gegs_atomic_probe.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, vp]
    L.gegs_ct_subsize.argtypes = [vp, vp, vp, vp, vp, vp]
    L.gegs_ct_resample_grid.argtypes = [vp, vp, vp, vp]
    L.gegs_ct_resample.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.gegs_set_tuning.argtypes = [vp, vp]
    L.gegs_run_batched.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
    L.gegs_load_mlc.argtypes = [vp, vp]
    L.gegs_load_jaws.argtypes = [vp, vp]
    L.gegs_load_block_wedge.argtypes = [vp, vp]
    L.gegs_set_fixes.argtypes = [vp, ctypes.c_uint32]
    L.gegs_set_jaw_positions.argtypes = [vp, vp, ctypes.c_uint32]
    L.gegs_set_leaf_positions.argtypes = [vp, vp, ctypes.c_uint32]
    L.gegs_set_kernel_timing.argtypes = [vp, ctypes.c_int]
    L.gegs_kernel_timing.argtypes = [vp, vp, vp]
    L.gegs_set_beamlets.argtypes = [vp, vp, ctypes.c_uint64, vp, ctypes.c_uint32]
    L.gegs_adjoint.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
    L.gegs_get_tuning.argtypes = [vp, vp]
    L.gegs_run_batched_device.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
    L.gegs_adjoint_device.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
    L.gegs_set_beamlet_weights_device.argtypes = [vp, vp, ctypes.c_uint32]
    for name in SYMBOLS:
        if getattr(L, name, None) is None:
            raise RuntimeError("symbol %s missing from %s" % (name, LIB_PATH))
    if L.gegs_sizeof_media_tables() != ctypes.sizeof(hatch.CMediaTables):
        raise RuntimeError("gegs_media_tables layout mismatch between include/gegs_media.h and hatch.py")
    _lib = L
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


FIX_Q1 = 1                 # GEGS_FIX_Q1: hownear_phantom from the particle's own voxel
FIX_Q8 = 2                 # GEGS_FIX_Q8: X-jaw step into the opening by the opening's distance
FIX_Q7 = 4                 # GEGS_FIX_Q7: range rejection on kinetic energy
FIX_FP64_FRAME = 8         # GEGS_FIX_FP64_FRAME: head -> phantom transform in binary64
FIX_Q2 = 16                # GEGS_FIX_Q2: pair bpar(1, medium) at 2 * medium
FIX_Q3 = 32                # GEGS_FIX_Q3: Compton Klein-Nishina resampling capped
FIX_ALL = 63


class GegsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gegs error %d: %s" % (code, msg))
        self.code = code


class Context:
    """One device context (gegs_ctx) with a workload loaded."""

    def __init__(self, device=0):
        self.L = load()
        self.h = self.L.gegs_create(device)
        if not self.h:
            raise RuntimeError("gegs_create failed: no usable HIP device %d" % device)
        self.nvox = 0
        self.nsrc = 0

    def _check(self, rc):
        if rc != 0:
            raise GegsError(rc, self.L.gegs_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.gegs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_media(self, tables):
        """The physics tables (hatch.MediaTables) of the media the regions use."""
        self._check(self.L.gegs_load_media(self.h, ctypes.c_void_p(tables.ptr())))

    def load_phantom(self, ph, regions):
        """The voxel phantom's bounds and its region records (phantom.region_table)."""
        reg = np.ascontiguousarray(regions)
        self._check(self.L.gegs_load_phantom(self.h, ph.nx, ph.ny, ph.nz, _p(ph.xb), _p(ph.yb), _p(ph.zb), _p(reg)))
        self.nvox = ph.nvox

    def load_mlc(self, mlc):
        """The MLC description (head.Mlc)."""
        self._check(self.L.gegs_load_mlc(self.h, ctypes.c_void_p(ctypes.addressof(mlc))))

    def load_jaws(self, jaws):
        """The secondary jaws (head.Jaws)."""
        self._check(self.L.gegs_load_jaws(self.h, ctypes.c_void_p(ctypes.addressof(jaws))))

    def load_workload(self, wl):
        self.load_media(wl.tables)
        self.load_phantom(wl.phantom, wl.regions)
        if getattr(wl, "mlc", None) is not None:
            self.load_mlc(wl.mlc)
            self.set_leaf_positions(wl.leaf_positions)
        if getattr(wl, "jaws", None) is not None:
            self.load_jaws(wl.jaws)
            self.set_jaw_positions(wl.jaw_positions)
        if getattr(wl, "block_wedge", None) is not None:
            self.load_block_wedge(wl.block_wedge)
        self._check(self.L.gegs_set_field(self.h, ctypes.c_void_p(ctypes.addressof(wl.field))))

    def set_source(self, src):
        src = np.ascontiguousarray(src)
        self._check(self.L.gegs_set_source(self.h, _p(src), len(src)))
        self.nsrc = len(src)

    def set_tuning(self, tuning=None, **fields):
        """Set the event-method tuning from a dict and/or keywords (zero or
        missing fields take the library defaults); None resets to defaults."""
        merged = dict(tuning or {})
        merged.update(fields)
        if not merged and tuning is None:
            self._check(self.L.gegs_set_tuning(self.h, None))
            return
        unknown = set(merged) - set(Tuning.FIELDS)
        if unknown:
            raise ValueError("unknown tuning fields: %s" % sorted(unknown))
        t = Tuning(**{k: int(v) for k, v in merged.items()})
        self._check(self.L.gegs_set_tuning(self.h, ctypes.byref(t)))

    def get_tuning(self):
        t = Tuning()
        self._check(self.L.gegs_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def atomic_probe(self, idx, nvox, reps=4):
        """FP64 atomic adds per second at the voxel indices idx (uint32 <
        nvox), `reps` timed passes (gegs_atomic_probe)."""
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        ms = ctypes.c_float()
        self._check(self.L.gegs_atomic_probe(self.h, _p(idx), len(idx), int(nvox), int(reps), ctypes.byref(ms)))
        return reps * len(idx) / (ms.value * 1e-3)

    def sample_events(self, particles, seed, max_out=16):
        """Run one event handler per particle (PARTICLE_DTYPE records);
        returns (out[n, max_out], n_out[n], edep[n])."""
        inp = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        n = len(inp)
        out = np.zeros(n * max_out, dtype=PARTICLE_DTYPE)
        n_out = np.zeros(n, dtype=np.uint32)
        edep = np.zeros(n, dtype=np.float64)
        self._check(self.L.gegs_sample_events(self.h, _p(inp), n, seed, _p(out), max_out, _p(n_out), _p(edep)))
        return out.reshape(n, max_out), n_out, edep

    @staticmethod
    def _jobs(jobs):
        arr = (Job * len(jobs))()
        for j, jb in enumerate(jobs):
            ctypes.memmove(ctypes.addressof(arr[j].field), ctypes.addressof(jb["field"]), 32)
            arr[j].plan = int(jb["plan"])
            arr[j].leaf_set = int(jb.get("leaf_set", 0))
            arr[j].first = int(jb["first"])
            arr[j].count = int(jb["count"])
            arr[j].history_offset = int(jb.get("history_offset", jb["first"]))
        return arr

    def set_leaf_positions(self, positions):
        """Leaf positions of one or more leaf sets (120 floats each: bank B
        +pos for leaves 0-59, then bank A); a job picks its set by index."""
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        if len(pos) == 0 or len(pos) % 120:
            raise ValueError("leaf positions come in sets of 120")
        self._check(self.L.gegs_set_leaf_positions(self.h, _p(pos), len(pos) // 120))

    def load_block_wedge(self, bw):
        """Block and wedge geometry (head.BlockWedge); has_block / has_wedge
        switch the modules on for this context."""
        self._check(self.L.gegs_load_block_wedge(self.h, ctypes.c_void_p(ctypes.addressof(bw))))

    def set_fixes(self, mask):
        """GEGS_FIX_* corrections of reference quirks for later runs of this
        context (0, the default: results as the reference's)."""
        self._check(self.L.gegs_set_fixes(self.h, int(mask)))

    def set_jaw_positions(self, xy):
        """Jaw openings of one or more sets, 4 floats each (x1, x2, y1, y2 in
        cm at the isocentre); set k goes with leaf set k."""
        a = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        if len(a) == 0 or len(a) % 4:
            raise ValueError("jaw positions come in sets of 4")
        self._check(self.L.gegs_set_jaw_positions(self.h, _p(a), len(a) // 4))

    def set_beamlets(self, beamlet=None, weight=None):
        """Beamlet index per history slot (job after job) and weight per
        beamlet; None removes the map."""
        self._beamlet_owner = None
        if beamlet is None:
            self._check(self.L.gegs_set_beamlets(self.h, None, 0, None, 0))
            return
        b = np.ascontiguousarray(beamlet, dtype=np.uint32)
        w = np.ascontiguousarray(weight, dtype=np.float32)
        self._check(self.L.gegs_set_beamlets(self.h, _p(b), len(b), _p(w), len(w)))

    def adjoint(self, jobs, n_plans, seed, dLdD, n_beamlets, n_batches=10):
        """Gradient of L with respect to the beamlet weights, given dL/dD of
        shape (n_plans, nvox) (replay of run_batched with the same seed)."""
        g = np.ascontiguousarray(dLdD, dtype=np.float64).reshape(-1)
        if len(g) != n_plans * self.nvox:
            raise ValueError("dLdD must have n_plans * nvox values")
        grad = np.zeros(n_beamlets, dtype=np.float64)
        opt = RunOptions(seed, 0, n_batches, 0, 0, 0, 0, 0)
        self._check(self.L.gegs_adjoint(self.h, self._jobs(jobs), len(jobs), n_plans, ctypes.byref(opt),
                                        _p(g), _p(grad)))
        return grad

    def set_beamlet_weights_device(self, weight_ptr, n_beamlets):
        """Weights of the current beamlet map from device memory (float32)."""
        self._check(self.L.gegs_set_beamlet_weights_device(self.h, ctypes.c_void_p(weight_ptr), n_beamlets))

    def run_batched_device(self, jobs, n_plans, seed, n_batches, dose_ptr, dose2_ptr=0):
        """run_batched writing dose (and dose2 unless 0) into device buffers
        of n_plans * nvox float64."""
        opt = RunOptions(seed, 0, n_batches, 0, 0, 0, 0, 0)
        self._check(self.L.gegs_set_history_tally(self.h, 0))
        self._check(self.L.gegs_run_batched_device(self.h, self._jobs(jobs), len(jobs), n_plans, ctypes.byref(opt),
                                                   ctypes.c_void_p(dose_ptr), ctypes.c_void_p(dose2_ptr or None)))

    def adjoint_device(self, jobs, n_plans, seed, dLdD_ptr, grad_ptr, n_batches=10):
        """adjoint with dL/dD (n_plans * nvox float64) and the gradient (one
        float64 per beamlet of the current map) in device memory."""
        opt = RunOptions(seed, 0, n_batches, 0, 0, 0, 0, 0)
        self._check(self.L.gegs_adjoint_device(self.h, self._jobs(jobs), len(jobs), n_plans, ctypes.byref(opt),
                                               ctypes.c_void_p(dLdD_ptr), ctypes.c_void_p(grad_ptr)))

    def run_batched(self, jobs, n_plans, seed, n_batches=10, per_history=False):
        """Event-queued transport of several fields/plans in one loop.

        jobs: sequence of dicts with keys field (workloads.Field), plan,
        first, count and optionally history_offset (default: first).
        Returns dose and dose2 of shape (n_plans, nvox)."""
        arr = self._jobs(jobs)
        self._check(self.L.gegs_set_history_tally(self.h, 1 if per_history else 0))
        opt = RunOptions(seed, 0, n_batches, 0, 0, 0, 0, 0)
        dose = np.zeros(n_plans * self.nvox, dtype=np.float64)
        dose2 = np.zeros(n_plans * self.nvox, dtype=np.float64)
        self._check(self.L.gegs_run_batched(self.h, arr, len(jobs), n_plans, ctypes.byref(opt), _p(dose), _p(dose2)))
        out = dict(dose=dose.reshape(n_plans, self.nvox), dose2=dose2.reshape(n_plans, self.nvox), n_batches=n_batches)
        ms = ctypes.c_float()
        self._check(self.L.gegs_last_kernel_ms(self.h, ctypes.byref(ms)))
        out["ms"] = ms.value
        cnt = Counters()
        self._check(self.L.gegs_get_counters(self.h, ctypes.byref(cnt)))
        out["counters"] = cnt.as_dict()
        if per_history:
            n = sum(int(jb["count"]) for jb in jobs)
            edep = np.zeros(n, dtype=np.float64)
            steps = np.zeros(n, dtype=np.uint32)
            self._check(self.L.gegs_history_tally(self.h, _p(edep), _p(steps)))
            out["edep"] = edep
            out["steps"] = steps
        return out

    def set_kernel_timing(self, enable=True):
        """Time every photon kernel launch of the event path with HIP events
        (resets the accumulators)."""
        self._check(self.L.gegs_set_kernel_timing(self.h, 1 if enable else 0))

    def kernel_timing(self):
        """(summed photon kernel launch durations in ms, launches) since the
        last set_kernel_timing."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        self._check(self.L.gegs_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def run_plain_noalloc(self, seed, dose_ptr, dose2_ptr, n_batches=10, history_offset=0, grid=0, depth=0,
                          first=0, count=0):
        """gegs_run_plain_noalloc: the plain kernel with no device allocation.
        dose_ptr, dose2_ptr: device addresses of nvox float64 each, written in
        place (the caller's writes to them must be complete).  Returns the
        run's counters."""
        self._check(self.L.gegs_set_history_tally(self.h, 0))
        opt = RunOptions(seed, history_offset, n_batches, 0, grid, depth, first, count)
        self._check(self.L.gegs_run_plain_noalloc(self.h, ctypes.c_void_p(ctypes.addressof(opt)),
                                                  ctypes.c_void_p(dose_ptr), ctypes.c_void_p(dose2_ptr)))
        cnt = Counters()
        self._check(self.L.gegs_get_counters(self.h, ctypes.byref(cnt)))
        return cnt.as_dict()

    def set_field(self, field):
        self._check(self.L.gegs_set_field(self.h, ctypes.c_void_p(ctypes.addressof(field))))

    def run(self, seed, n_batches=10, history_offset=0, method="plain", per_history=False,
            grid=0, block=0, depth=0, first=0, count=0):
        self._check(self.L.gegs_set_history_tally(self.h, 1 if per_history else 0))
        opt = RunOptions(seed, history_offset, n_batches, block, grid, depth, first, count)
        dose = np.zeros(self.nvox, dtype=np.float64)
        dose2 = np.zeros(self.nvox, dtype=np.float64)
        fn = self.L.gegs_run_plain if method == "plain" else self.L.gegs_run_event
        self._check(fn(self.h, ctypes.c_void_p(ctypes.addressof(opt)), _p(dose), _p(dose2)))
        out = dict(dose=dose, dose2=dose2, n_batches=n_batches)
        ms = ctypes.c_float()
        self._check(self.L.gegs_last_kernel_ms(self.h, ctypes.byref(ms)))
        out["ms"] = ms.value
        cnt = Counters()
        self._check(self.L.gegs_get_counters(self.h, ctypes.byref(cnt)))
        out["counters"] = cnt.as_dict()
        if per_history:
            edep = np.zeros(self.nsrc, dtype=np.float64)
            steps = np.zeros(self.nsrc, dtype=np.uint32)
            self._check(self.L.gegs_history_tally(self.h, _p(edep), _p(steps)))
            out["edep"] = edep
            out["steps"] = steps
        return out
