"""Seeded synthetic workloads of the benchmark (phantom + media + field + source).

``water``      60^3 water at 0.5 cm, parallel 10x10 cm^2 field of 1.25 MeV
               photons on the +z face, ECUT 0.7 MeV total / PCUT 0.01 MeV.
``ct``         128x128x100 synthetic CT at 0.25 cm through the ctcreate ramp
               (air, lung, ICRU tissue, ICRP bone); point source at SSD 100 cm,
               10x10 cm^2 field, 24-bin 6 MV-like spectrum; CT axis convention
               (the beam travels along phantom +y).
"""
from __future__ import annotations

import ctypes
import functools
import math

import numpy as np

from . import hatch, phantom as ph, source as src

M_PHANTOM = 6


class Field(ctypes.Structure):
    """ctypes mirror of ``gegs_field`` (include/gegs.h)."""
    _fields_ = [("change_xyz", ctypes.c_int32), ("module_first", ctypes.c_int32),
                ("source_to_isocenter", ctypes.c_float), ("cos_angle", ctypes.c_float),
                ("sin_angle", ctypes.c_float), ("iso_x", ctypes.c_float),
                ("iso_y", ctypes.c_float), ("iso_z", ctypes.c_float)]


def make_field(change_xyz, sid, gantry_deg=0.0, iso=(0.0, 0.0, 0.0), module_first=M_PHANTOM):
    a = np.float32(gantry_deg * 3.1415926 / 180.0)     # CUDA_EGS.cu:424
    # cos and sin of the binary32 angle in binary64, rounded once (the C
    # library's, as the C++ run driver evaluates them)
    return Field(int(change_xyz), int(module_first), float(sid), float(np.float32(math.cos(float(a)))),
                 float(np.float32(math.sin(float(a)))), *map(float, iso))


class Workload:
    def __init__(self, name, phantom, media_names, field, other_media, source_fn, mlc=None, leaf_positions=None,
                 mlc_medium=None, jaws=None, jaw_positions=None):
        self.name = name
        self.jaws = jaws                      # head.Jaws, fields starting in the jaws
        self.jaw_positions = jaw_positions    # x1, x2, y1, y2 of the field
        self.mlc = mlc                        # head.Mlc, fields starting in the MLC
        self.leaf_positions = leaf_positions  # 120 leaf positions of the field
        self.mlc_medium = mlc_medium
        self.phantom = phantom
        self.media_names = media_names
        self.field = field
        self.other_media = other_media
        self._source_fn = source_fn
        self._tables = None
        self._regions = None
        self.block_wedge = None               # head.BlockWedge, fields passing a block or a wedge

    @property
    def tables(self):
        if self._tables is None:
            self._tables = _tables(tuple(self.media_names), self.mlc_medium)
        return self._tables

    @property
    def regions(self):
        if self._regions is None:
            self._regions = ph.region_table(self.phantom, self.tables, self.other_media)
        return self._regions

    def source(self, n, seed):
        return self._source_fn(n, seed)


@functools.lru_cache(maxsize=8)
def _tables(names, mlc_medium=None):
    return hatch.MediaTables(list(names), mlc_medium=mlc_medium)


def water(n=60, voxel=0.5, energy=1.25, field_half=5.0):
    phan = ph.water_phantom(n, n * voxel)
    fld = make_field(False, 100.0)
    return Workload("water_60cube_1.25MeV", phan, ["H2O700ICRU", "AIR700ICRU"], fld,
                    ["AIR700ICRU"] * 5,
                    lambda k, seed: src.parallel_beam(k, energy, field_half, 90.0, seed))


def ct(nx=128, ny=128, nz=100, voxel=0.25, seed=2):
    phan = ph.ct_phantom(nx, ny, nz, voxel, seed)
    body_half_ap = 0.30 * ny * voxel
    sid = 100.0 + body_half_ap
    fld = make_field(True, sid)
    names = [r[0] for r in ph.DEFAULT_RAMP]
    return Workload("ct_128x128x100_6MV", phan, names, fld, ["AIR700ICRU"] * 5,
                    lambda k, s: src.point_source_spectrum(k, 100.0, 10.0, s))


PHSP_Z = 16.44          # phase-space plane of the reference's run inputs (Test*.ini, [Phase space])


def ct_mlc(gantry_deg=0.0, aperture_seed=3, nx=128, ny=128, nz=100, voxel=0.25, seed=2):
    """The CT workload behind a 60-pair MLC: a synthetic head phase space at
    z = 16.44 cm (3% electrons) enters the MLC (synthetic Millennium-like
    geometry, seeded irregular aperture inside 10 x 10 cm^2), then the
    phantom.  Regions 1-5: air, jaw W, MLC W, air, air."""
    from . import head
    phan = ph.ct_phantom(nx, ny, nz, voxel, seed)
    sid = 100.0 + 0.30 * ny * voxel
    fld = make_field(True, sid, gantry_deg, module_first=head.M_MLC)
    names = [r[0] for r in ph.DEFAULT_RAMP] + ["W700ICRU"]
    mlc = head.build_mlc(head.millennium_like())
    pos = head.random_aperture(aperture_seed)
    return Workload("ct_128x128x100_6MV_mlc", phan, names, fld,
                    ["AIR700ICRU", "W700ICRU", "W700ICRU", "AIR700ICRU", "AIR700ICRU"],
                    lambda k, s: src.head_phase_space(k, PHSP_Z, 100.0, 10.0, s),
                    mlc=mlc, leaf_positions=pos, mlc_medium="W700ICRU")


def ct_jaws_mlc(gantry_deg=0.0, aperture_seed=3, seed=2):
    """The MLC workload with the secondary jaws above the MLC (Clinac-like
    slabs at z 28-35.8 cm and 36.7-44.5 cm) framing the aperture: fields
    start in the Y jaws, pass the X jaws and the MLC, then the phantom.
    Regions 1-5: air, jaw W, MLC W, air, air."""
    from . import head
    wl = ct_mlc(gantry_deg, aperture_seed, seed=seed)
    wl.name = "ct_128x128x100_6MV_jaws_mlc"
    wl.field = make_field(True, wl.field.source_to_isocenter, gantry_deg, module_first=head.M_SECJAWY)
    wl.jaws = head.build_jaws(**head.CLINAC_LIKE_JAWS)
    wl.jaw_positions = head.jaw_positions_for(wl.leaf_positions)
    return wl


# A shielding block and a physical wedge below the jaws (head frame, cm from
# the source), as BEAMnrc PYRAMIDS values.  The block: 3 cm of tungsten at
# z 55-58 cm with source-focused faces; its "X" values (which, as the
# reference builds, bound the head's y) cover y in [0.55, 2.2] cm at the
# block's top, its "Y" values x in [-1.65, 1.65] cm.  The wedge: three
# 0.5 cm layers at z 60-61.5 cm whose - side steps from 2.4 to -1.2 cm, a
# tungsten wedge with a 1.5 cm heel, over |x| <= 10 cm.
BLOCK_PYRAMIDS = [55.0, 58.0, 2.2, 2.32, 0.55, 0.58, 1.65, 1.74, -1.65, -1.74]
WEDGE_PYRAMIDS = (10.0, [[60.0, 60.5, 2.4, 2.4, 2.4, 1.2],
                         [60.5, 61.0, 2.4, 2.4, 1.2, 0.0],
                         [61.0, 61.5, 2.4, 2.4, 0.0, -1.2]])


def ct_block_wedge(module_first=None, block=True, wedge=True, gantry_deg=0.0, seed=2):
    """The CT workload (128x128x100 at 2.5 mm) under a block and/or a wedge:
    the synthetic head phase space at z = 16.44 cm, aimed at x in [-2.5, 2.5],
    y in [0, 5] cm at 100 cm (over the block and the wedge, and partly beside
    them), enters the first loaded of block, wedge, then the phantom.
    Regions 1-5: air, jaw W, MLC W, block W, wedge W."""
    from . import head
    phan = ph.ct_phantom(128, 128, 100, 0.25, seed)
    sid = 100.0 + 0.30 * 128 * 0.25
    if module_first is None:
        module_first = head.M_BLOCK if block else head.M_WEDGE
    fld = make_field(True, sid, gantry_deg, module_first=module_first)
    names = [r[0] for r in ph.DEFAULT_RAMP] + ["W700ICRU"]
    bw = head.build_block_wedge(BLOCK_PYRAMIDS if block else None, WEDGE_PYRAMIDS if wedge else None)
    wl = Workload("ct_128x128x100_6MV_" + "_".join(n for n, on in (("block", block), ("wedge", wedge)) if on),
                  phan, names, fld, ["AIR700ICRU", "W700ICRU", "W700ICRU", "W700ICRU", "W700ICRU"],
                  lambda k, s: src.head_phase_space(k, PHSP_Z, 100.0, 3.0, s, center=(0.0, 2.5)))
    wl.block_wedge = bw
    return wl


C3_GANTRY = (0.0, 72.0, 144.0, 216.0, 288.0)


def field_weights(n, lo, hi, seed):
    """n field weights in [lo, hi] that sum to 1 (seeded; redrawn until the
    normalised draw stays inside the interval)."""
    rng = np.random.default_rng(seed)
    for _ in range(100000):
        w = rng.uniform(lo, hi, n)
        w = w / w.sum()
        if w.min() >= lo and w.max() <= hi:
            return w
    raise ValueError("no weights in [%g, %g] sum to 1 for n = %d" % (lo, hi, n))


class Plan:
    """Fields of one plan sharing a phantom, an MLC and one phase-space list:
    field j uses leaf set j and the records [offset_j, offset_j + count_j)."""

    def __init__(self, name, workload, fields, leaf_positions, weights, source_fn):
        self.name = name
        self.workload = workload
        self.fields = fields
        self.leaf_positions = np.asarray(leaf_positions, dtype=np.float32)
        self.weights = np.asarray(weights, dtype=np.float64)
        self._source = source_fn

    def source(self, n, seed):
        return self._source(n, seed)

    def counts(self, n_total):
        """Histories per field: the weights times n_total, rounded, summing to n_total."""
        c = np.floor(self.weights * n_total).astype(np.int64)
        c[np.argmax(self.weights)] += n_total - int(c.sum())
        return c

    def jobs(self, n_total, part=0, n_parts=1, plan=0):
        """Jobs of part `part` of n_parts equal slices of every field (for
        gegs_run_batched with the plan's leaf sets loaded in field order)."""
        c = self.counts(n_total)
        off = np.concatenate([[0], np.cumsum(c)[:-1]])
        out = []
        for j, f in enumerate(self.fields):
            lo = int(c[j] * part // n_parts)
            hi = int(c[j] * (part + 1) // n_parts)
            if hi > lo:
                out.append(dict(field=f, plan=plan, first=int(off[j]) + lo, count=hi - lo, leaf_set=j))
        return out


def mlc_plan(gantry=C3_GANTRY, seed=3, weight_range=(0.1, 0.35)):
    """Five MLC fields on the CT phantom: gantry 0/72/144/216/288 degrees,
    seeded irregular apertures inside 10 x 10 cm^2, uneven field weights in
    [0.1, 0.35], one egsphsp MODE0 phase-space list (decoded) shared by the
    fields in disjoint slices."""
    from . import head
    wl = ct_mlc()
    sid = wl.field.source_to_isocenter
    fields = [make_field(True, sid, g, module_first=head.M_MLC) for g in gantry]
    pos = np.stack([head.random_aperture(seed * 1000 + j) for j in range(len(fields))])
    w = field_weights(len(fields), weight_range[0], weight_range[1], seed)

    def phsp(n, s):
        rec = src.encode_phsp(src.head_phase_space(n, PHSP_Z, 100.0, 10.0, s))
        return src.decode_phsp(rec, PHSP_Z)

    return Plan("ct_mlc_5field", wl, fields, pos, w, phsp)


def ct_cube(n=128, voxel=0.25, seed=4):
    """The dose-optimisation phantom: the synthetic CT at n^3 voxels of
    `voxel` cm (128^3 x 2.5 mm), point source at SSD 100 with a 10 x 10 field."""
    wl = ct(nx=n, ny=n, nz=n, voxel=voxel, seed=seed)
    wl.name = "ct_%dcube_6MV" % n
    return wl


def target_and_oar(phantom, r_target=3.0, oar=(4.0, 5.0)):
    """Voxel masks of the optimisation objective: a target sphere of radius
    r_target cm at the isocentre (the phantom centre) and an organ-at-risk
    shell oar[0] <= r < oar[1] cm around it."""
    xc = 0.5 * (phantom.xb[:-1] + phantom.xb[1:]).astype(np.float64)
    yc = 0.5 * (phantom.yb[:-1] + phantom.yb[1:]).astype(np.float64)
    zc = 0.5 * (phantom.zb[:-1] + phantom.zb[1:]).astype(np.float64)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).reshape(-1)
    return r < r_target, (r >= oar[0]) & (r < oar[1])


def objective(phantom, dose, boost=1.2):
    """L = sum_v m_v (D_v - P_v)^2 with m = 1 on the target and the OAR shell,
    P = boost x the mean target dose of `dose` on the target and 0 on the OAR.
    Returns (m, P); dL/dD = 2 m (D - P)."""
    target, oar = target_and_oar(phantom)
    m = (target | oar).astype(np.float64)
    p = np.where(target, boost * float(dose[target].mean()), 0.0)
    return m, p


class MultiPlan:
    """Several plans on one phantom, each with its own fields and history
    budget, all fed by one point-source record list: field j of plan k uses
    records [first, first + count) of that list, in plan-then-field order."""

    def __init__(self, name, workload, plans, budgets):
        self.name = name
        self.workload = workload
        self.plans = plans              # list of lists of Field
        self.budgets = np.asarray(budgets, dtype=np.int64)

    @property
    def n_histories(self):
        return int(self.budgets.sum())

    def source(self, seed):
        return self.workload.source(self.n_histories, seed)

    def jobs(self):
        out, first = [], 0
        for k, fields in enumerate(self.plans):
            b = int(self.budgets[k])
            for j, f in enumerate(fields):
                c = b * (j + 1) // len(fields) - b * j // len(fields)
                out.append(dict(field=f, plan=k, first=first, count=c))
                first += c
        return out


def multi_plan(n_plans=16, total=100_000_000, n=(256, 256, 180), voxel=0.2, seed=5,
               fields=(1, 9), budget=(1_000_000, 20_000_000)):
    """The plan-queue workload: the synthetic CT at 256 x 256 x 180 voxels of
    2 mm, n_plans plans of 1-9 coplanar fields at seeded gantry angles, history
    budgets in [1e6, 2e7] summing to `total` (1e8); 6 MV-like point source,
    SSD 100, 10 x 10 cm^2 fields."""
    rng = np.random.default_rng(seed)
    wl = ct(nx=n[0], ny=n[1], nz=n[2], voxel=voxel, seed=seed)
    wl.name = "ct_%dx%dx%d_6MV" % n
    sid = wl.field.source_to_isocenter
    plans = []
    for _ in range(n_plans):
        k = int(rng.integers(fields[0], fields[1] + 1))
        g0 = float(rng.uniform(0.0, 360.0))
        plans.append([make_field(True, sid, (g0 + 360.0 * j / k) % 360.0) for j in range(k)])
    lo, hi = budget
    for _ in range(100000):
        b = rng.uniform(lo, hi, n_plans)
        b = np.floor(b / b.sum() * total).astype(np.int64)
        b[np.argmax(b)] += total - int(b.sum())
        if b.min() >= lo and b.max() <= hi:
            break
    else:
        raise ValueError("no budgets in [%g, %g] sum to %d" % (lo, hi, total))
    return MultiPlan("ct_%dx%dx%d_%dplans" % (n[0], n[1], n[2], n_plans), wl, plans, b)
