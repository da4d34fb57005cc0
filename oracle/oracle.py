"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", HERE, "-B" if force else "all"])
    return LIB


def build_native(dest_dir):
    """The oracle built for the CPU of this host (gcc -O3 -march=native) into
    dest_dir; returns its path."""
    out = os.path.join(dest_dir, "liboracle_native.so")
    subprocess.check_call(["make", "-s", "-C", HERE, "-B", "native", "NATIVE_OUT=" + out])
    return out


def use_library(path):
    """Load the oracle from `path` from now on (None: oracle/liboracle.so)."""
    global _lib, _path
    _lib = None
    _path = path


class Counters(ctypes.Structure):
    _fields_ = [("photon_steps", ctypes.c_uint64), ("electron_steps", ctypes.c_uint64),
                ("region_gathers", ctypes.c_uint64), ("deposits", ctypes.c_uint64),
                ("interactions", ctypes.c_uint64 * 10), ("histories", ctypes.c_uint64),
                ("overflow", ctypes.c_uint64)]


_lib = None
_path = None


def lib():
    global _lib
    if _lib is None:
        if _path is None:
            build()
        L = ctypes.CDLL(_path or LIB)
        vp = ctypes.c_void_p
        L.oracle_run.restype = ctypes.c_int
        L.oracle_run.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 vp, vp, vp, vp, vp, ctypes.c_uint64, ctypes.c_uint64,
                                 ctypes.c_uint64, vp, vp, vp, vp, vp, vp]
        L.oracle_run_mt.restype = ctypes.c_int
        L.oracle_run_mt.argtypes = L.oracle_run.argtypes + [ctypes.c_int, vp, vp, vp, vp]
        L.oracle_jaw_polys.argtypes = [vp, vp, ctypes.c_float, vp]
        L.oracle_howfar_jaws.restype = ctypes.c_uint32
        L.oracle_howfar_jaws.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.oracle_set_fixes.argtypes = [ctypes.c_uint32]
        L.oracle_howfar_block_wedge.restype = ctypes.c_uint32
        L.oracle_howfar_block_wedge.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.oracle_leaf_ends.argtypes = [vp, vp, ctypes.c_float, vp]
        L.oracle_leaf_index.restype = ctypes.c_int
        L.oracle_leaf_index.argtypes = [vp, ctypes.c_float, ctypes.c_float]
        L.oracle_howfar_mlc.restype = ctypes.c_uint32
        L.oracle_howfar_mlc.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.oracle_where_mlc.restype = ctypes.c_uint32
        L.oracle_where_mlc.argtypes = [vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.oracle_interaction.restype = ctypes.c_int
        L.oracle_interaction.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                         vp, vp, vp, vp, vp, ctypes.c_uint64, vp, ctypes.c_int, vp]
        L.oracle_mscat.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                                   ctypes.c_uint64, ctypes.c_uint32, vp]
        L.oracle_last_path.restype = ctypes.c_double
        L.oracle_relax.restype = ctypes.c_int
        L.oracle_relax.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                                   ctypes.c_uint64, vp, ctypes.c_int, vp]
        L.oracle_eii_shell.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.c_uint32, vp, vp]
        L.oracle_msdist.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, vp]
        L.oracle_compute_eloss.restype = ctypes.c_float
        L.oracle_compute_eloss.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_float]
        L.oracle_range.restype = ctypes.c_float
        L.oracle_range.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float]
        L.oracle_logf.restype = ctypes.c_float
        L.oracle_logf.argtypes = [ctypes.c_float]
        L.oracle_expf.restype = ctypes.c_float
        L.oracle_expf.argtypes = [ctypes.c_float]
        L.oracle_sincos2pi.argtypes = [ctypes.c_float, vp, vp]
        L.oracle_philox.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_primary_lineage.restype = ctypes.c_uint64
        L.oracle_primary_lineage.argtypes = [ctypes.c_uint64]
        L.oracle_sizeof_media_tables.restype = ctypes.c_size_t
        L.oracle_howfar_phantom.restype = ctypes.c_uint32
        L.oracle_howfar_phantom.argtypes = [ctypes.c_uint32] * 3 + [vp, vp, vp, vp, vp, ctypes.c_uint32, vp]
        L.oracle_to_phantom.argtypes = [vp, vp, vp]
        L.oracle_hownear_phantom.restype = ctypes.c_float
        L.oracle_hownear_phantom.argtypes = [ctypes.c_uint32] * 3 + [vp, vp, vp, vp, vp, ctypes.c_uint32]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def run(workload, source, seed, hist_offset=0, per_history=False, threads=1):
    """Transport ``source`` (SOURCE_DTYPE array) through the workload.
    Returns dict(dose, dose2, edep, steps, counters)."""
    L = lib()
    phan = workload.phantom
    xb, yb, zb = phan.padded_bounds()
    n = len(source)
    src = np.ascontiguousarray(source)
    dose = np.zeros(phan.nvox, dtype=np.float64)
    dose2 = np.zeros(phan.nvox, dtype=np.float64)
    edep = np.zeros(n, dtype=np.float64) if per_history else None
    steps = np.zeros(n, dtype=np.uint32) if per_history else None
    cnt = Counters()
    bal = np.zeros(2, dtype=np.float64)
    reg = np.ascontiguousarray(workload.regions)
    mlc_ptr, ends_ptr, ends, jaws = None, None, None, None
    first = workload.field.module_first
    sid = workload.field.source_to_isocenter
    mlc = getattr(workload, "mlc", None)
    if first == 3 or (first == 1 and mlc is not None):    # the field passes the MLC
        mlc_ptr = ctypes.c_void_p(ctypes.addressof(mlc))
        ends = leaf_ends(mlc, workload.leaf_positions, sid)
        ends_ptr = _p(ends)
    if first == 1:                                        # the field starts in the jaws
        jaws = jaw_polys(workload.jaws, workload.jaw_positions, sid)
    bw = getattr(workload, "block_wedge", None)
    bw_ptr = ctypes.c_void_p(ctypes.addressof(bw)) if bw is not None else None
    rc = L.oracle_run_mt(ctypes.c_void_p(workload.tables.ptr()), _p(reg), phan.nx, phan.ny, phan.nz,
                         _p(xb), _p(yb), _p(zb), ctypes.c_void_p(ctypes.addressof(workload.field)),
                         _p(src), n, hist_offset, seed, _p(dose), _p(dose2), _p(edep), _p(steps),
                         ctypes.c_void_p(ctypes.addressof(cnt)), _p(bal), int(threads), mlc_ptr, ends_ptr,
                         _p(jaws), bw_ptr)
    if rc != 0:
        raise RuntimeError("oracle_run failed: %d" % rc)
    return dict(dose=dose, dose2=dose2, edep=edep, steps=steps, counters=cnt,
                escaped=float(bal[0]), dropped=float(bal[1]))


def jaw_polys(jaws, xy, sid):
    """[32, 2] float32 jaw polygons of one field (x1, x2, y1, y2 at the isocentre)."""
    xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(4)
    out = np.zeros((32, 2), dtype=np.float32)
    lib().oracle_jaw_polys(ctypes.c_void_p(ctypes.addressof(jaws)), _p(xy), float(sid), _p(out))
    return out


def howfar_jaws(polys, is_y, jawx_next, particle, t):
    """(region, t, module) of the oracle's howfar_SecJawsY / X."""
    part = np.ascontiguousarray(particle)
    tt = ctypes.c_float(t)
    mod = ctypes.c_int32(0)
    r = lib().oracle_howfar_jaws(_p(np.ascontiguousarray(polys, dtype=np.float32)), int(is_y), int(jawx_next),
                                 _p(part), ctypes.byref(mod), ctypes.byref(tt))
    return int(r), tt.value, mod.value


def set_fixes(mask):
    """GEGS_FIX_* corrections of reference quirks for every later oracle call
    (0: the reference's behaviour)."""
    lib().oracle_set_fixes(int(mask))


def howfar_block_wedge(bw, is_wedge, particle, t):
    """(region, t, module, tperp) of the oracle's howfar_block / howfar_wedge
    and hownear for one particle."""
    part = np.ascontiguousarray(particle)
    tt = ctypes.c_float(t)
    tp = ctypes.c_float(0.0)
    mod = ctypes.c_int32(0)
    r = lib().oracle_howfar_block_wedge(ctypes.c_void_p(ctypes.addressof(bw)), int(is_wedge), _p(part),
                                        ctypes.byref(mod), ctypes.byref(tt), ctypes.byref(tp))
    return int(r), tt.value, mod.value, tp.value


def leaf_ends(mlc, positions, sid):
    """[60, 34, 2] float32 leaf ends of one leaf set (120 positions)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    out = np.zeros((60, 34, 2), dtype=np.float32)
    lib().oracle_leaf_ends(ctypes.c_void_p(ctypes.addressof(mlc)), _p(pos), float(sid), _p(out))
    return out


def leaf_index(mlc, x, z):
    return lib().oracle_leaf_index(ctypes.c_void_p(ctypes.addressof(mlc)), float(x), float(z))


def where_mlc(mlc, ends, x, y, z):
    return lib().oracle_where_mlc(ctypes.c_void_p(ctypes.addressof(mlc)), _p(ends), float(x), float(y), float(z))


def howfar_mlc(mlc, ends, particle, leaf, bank, t):
    """(region, t, leaf, bank, module) of the oracle's howfar_MLC."""
    part = np.ascontiguousarray(np.array([particle]).astype(particle.dtype))
    lf = ctypes.c_int32(leaf)
    bk = ctypes.c_int32(bank)
    md = ctypes.c_int32(0)
    tt = ctypes.c_float(t)
    r = lib().oracle_howfar_mlc(ctypes.c_void_p(ctypes.addressof(mlc)), _p(ends), _p(part), ctypes.byref(lf),
                                ctypes.byref(bk), ctypes.byref(md), ctypes.byref(tt))
    return r, tt.value, lf.value, bk.value, md.value


def mscat(tables, lam, chia2, q1, n, seed=1, lin0=1):
    """n multiple-scattering polar cosines from the oracle's mscat."""
    out = np.zeros(n, dtype=np.float32)
    lib().oracle_mscat(ctypes.c_void_p(tables.ptr()), lam, chia2, q1, seed, lin0, n, _p(out))
    return out


def interaction(workload, particle, seed, max_out=64):
    """Run the handler of particle['status'] once (PARTICLE_DTYPE record from
    gegs_amd.native); returns (pushed records, scored energy)."""
    from gegs_amd.native import PARTICLE_DTYPE
    L = lib()
    phan = workload.phantom
    xb, yb, zb = phan.padded_bounds()
    inp = np.ascontiguousarray(np.asarray(particle, dtype=PARTICLE_DTYPE).reshape(1))
    out = np.zeros(max_out, dtype=PARTICLE_DTYPE)
    ed = np.zeros(1, dtype=np.float64)
    reg = np.ascontiguousarray(workload.regions)
    k = L.oracle_interaction(ctypes.c_void_p(workload.tables.ptr()), _p(reg), phan.nx, phan.ny, phan.nz,
                             _p(xb), _p(yb), _p(zb), ctypes.c_void_p(ctypes.addressof(workload.field)),
                             _p(inp), seed, _p(out), max_out, _p(ed))
    return out[:k].copy(), float(ed[0])


def relax(tables, region, iz, shell, energy, seed, lin, max_out=32):
    """The relaxation cascade of one vacancy: (particles created, local deposit)."""
    from gegs_amd.native import PARTICLE_DTYPE
    out = np.zeros(max_out, dtype=PARTICLE_DTYPE)
    ed = np.zeros(1, dtype=np.float32)
    reg = np.ascontiguousarray(np.asarray(region).reshape(1))
    k = lib().oracle_relax(ctypes.c_void_p(tables.ptr()), _p(reg), iz, shell, energy, seed, lin, _p(out), max_out,
                           _p(ed))
    return out[:k].copy(), float(ed[0])


def eii_shell(tables, medium, ekin, n, seed=1, lin0=1):
    """(shell, Z) chosen by the EII branch of moller for n electrons (0: Moller)."""
    sh = np.zeros(n, dtype=np.int32)
    z = np.zeros(n, dtype=np.int32)
    lib().oracle_eii_shell(ctypes.c_void_p(tables.ptr()), medium, ekin, seed, lin0, n, _p(sh), _p(z))
    return sh, z


def last_path():
    """Electron path length (sum of step lengths) of the last interaction() call."""
    return float(lib().oracle_last_path())


def msdist(tables, medium, rhof, e, eloss, tustep, direction, n, seed=1, lin0=1):
    """n PRESTA-I steps (msdist_pI) from the origin: array n x 7 of displacement
    (x, y, z), final direction (u, v, w) and the returned straight-line distance."""
    out = np.zeros((n, 7), dtype=np.float32)
    d = np.ascontiguousarray(direction, dtype=np.float32)
    lib().oracle_msdist(ctypes.c_void_p(tables.ptr()), medium, rhof, e, eloss, tustep, _p(d), seed, lin0, n, _p(out))
    return out


def compute_eloss(tables, qel, medium, range_, rhof, step, eke):
    return float(lib().oracle_compute_eloss(ctypes.c_void_p(tables.ptr()), qel, medium, range_, rhof, step, eke))


def electron_range(tables, qel, medium, rhof, eke):
    return float(lib().oracle_range(ctypes.c_void_p(tables.ptr()), qel, medium, rhof, eke))


def to_phantom(field, pos_dir):
    """Head -> phantom frame of points and directions (N x 6 float32)."""
    L = lib()
    pd = np.ascontiguousarray(pos_dir, dtype=np.float32).reshape(-1, 6)
    out = np.zeros_like(pd)
    for i in range(len(pd)):
        L.oracle_to_phantom(ctypes.byref(field), _p(pd[i]), _p(out[i]))
    return out
