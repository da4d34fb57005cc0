"""Treatment-head inputs: the DYNVMLC multileaf collimator and its fields.

The reference reads the MLC from the DYNVMLC block of a BEAMnrc .egsinp file
(read_egsinp, init.cu:345-611): leaf dimensions of the three leaf types
(full, target, isocentre), the number of leaves, the leaf row start and air
gap, the leaf-end radius and the default openings.  It builds the leaf
cross-sections across the leaves (Leaf::VerCal_*, read_egsinp.h:37-86), with
sides focused on the source, the closed leaf-end profile along the leaf
travel (leaf_end, init.cu:256-343) and, per field, the leaf ends shifted to
the field's leaf positions (init_leaf_pos, init.cu:1594-1662).  Fields come
from a .mlcinfo list (read_mlcinp, init.cu:1377-1500): gantry angle, jaws,
MU and a Varian leaf file of 120 positions.

The reference's VarianMLC.egsinp and .mlcinfo files are not available, so
``MILLENNIUM_LIKE`` is a synthetic 120-leaf geometry of this build (40 inner
pairs of 5 mm and 20 outer pairs of 10 mm projected at the isocentre; leaf
height, tongue-and-groove and rail dimensions chosen to be physically
plausible, not taken from a vendor drawing).  The leaf-end profile is a
circular arc of the leaf-end radius sampled at equal angles (the reference
hard-codes the 15 points of its arc).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

NLEAF = 60
M_SECJAWY = 1
M_SECJAWX = 2
M_MLC = 3


class Float2(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float)]


class Jaws(ctypes.Structure):
    """ctypes mirror of ``gegs_jaws`` (include/gegs.h): the JAWS component of
    the head (init.cu:612-688), z in cm from the source."""
    _fields_ = [("rmax", ctypes.c_float), ("zminy", ctypes.c_float), ("zmaxy", ctypes.c_float),
                ("zminx", ctypes.c_float), ("zmaxx", ctypes.c_float)]


# jaw slabs of a Clinac-type head (z of the Y and X pairs, cm from the source)
CLINAC_LIKE_JAWS = dict(rmax=20.0, zminy=28.0, zmaxy=35.8, zminx=36.7, zmaxx=44.5)


def build_jaws(rmax, zminy, zmaxy, zminx, zmaxx):
    return Jaws(rmax, zminy, zmaxy, zminx, zmaxx)


def jaw_positions_for(leaf_positions, margin=0.5, half=5.0):
    """Jaw openings x1, x2, y1, y2 (cm at the isocentre) that frame an MLC
    aperture: across the leaves the open pairs' extent, along the leaves the
    leaf ends' extent, each widened by `margin` and kept within +-(half +
    margin)."""
    pos = np.asarray(leaf_positions, dtype=np.float64).reshape(2 * NLEAF)
    open_ = pos[:NLEAF] + pos[NLEAF:] > 0
    centers = np.concatenate([np.arange(10) * 1.0 - 19.5, np.arange(40) * 0.5 - 9.75, np.arange(10) * 1.0 + 10.5])
    width = np.concatenate([np.full(10, 1.0), np.full(40, 0.5), np.full(10, 1.0)])
    lim = half + margin
    if not open_.any():
        return np.array([-margin, margin, -margin, margin], dtype=np.float32)
    y1 = max(-lim, float((centers - width / 2)[open_].min()) - margin)
    y2 = min(lim, float((centers + width / 2)[open_].max()) + margin)
    x1 = max(-lim, float(-pos[NLEAF:][open_].max()) - margin)
    x2 = min(lim, float(pos[:NLEAF][open_].max()) + margin)
    return np.array([x1, x2, y1, y2], dtype=np.float32)


class Mlc(ctypes.Structure):
    """ctypes mirror of ``gegs_mlc`` (include/gegs.h)."""
    _fields_ = [("zmin", ctypes.c_float), ("zthick", ctypes.c_float), ("rmax", ctypes.c_float),
                ("x_start", ctypes.c_float), ("x_end", ctypes.c_float), ("x_mid_1", ctypes.c_float),
                ("x_mid_2", ctypes.c_float), ("air_gap", ctypes.c_float), ("wl_full", ctypes.c_float),
                ("wl_target", ctypes.c_float), ("wl_isocenter", ctypes.c_float), ("wt", ctypes.c_float),
                ("cross_xz", Float2 * (NLEAF * 16)), ("boundary_xz", Float2 * 4), ("boundary_yz", Float2 * 4),
                ("cross_yz", Float2 * 34)]

    def cross(self):
        return np.ctypeslib.as_array(self.cross_xz).view(np.float32).reshape(NLEAF, 16, 2)


@dataclasses.dataclass
class Leaf:
    """DYNVMLC leaf type: widths (wl leaf, wt tongue, wg groove, wtip tip,
    wts/wbs top/bottom rail) and z planes (absolute, from the source)."""
    wl: float
    wt: float
    wg: float
    wtip: float
    wts: float
    wbs: float
    ztip: float
    zl: float
    zt: float
    zg: float
    zth: float
    zbh: float
    holepos: float
    zts: float
    zbs: float


@dataclasses.dataclass
class Dynvmlc:
    rmax: float
    zmin: float
    zthick: float
    full: Leaf
    target: Leaf
    isocenter: Leaf
    n_full: int            # full leaves on each side
    n_half: int            # half leaves in the middle (target + isocentre)
    x_start: float
    air_gap: float
    rad_end: float
    openings: list         # [(neg, pos, count), ...] covering 60 leaves
    medium: str = "W700ICRU"


def _vertices(kind, lf, zmin, zthick):
    """The 14 cross-section vertices (x, z) of a leaf (read_egsinp.h:37-86)."""
    zmax = zmin + zthick
    if kind == "full":
        v = [(0, lf.ztip), (lf.wtip, lf.ztip), (lf.wtip, lf.zl), (lf.wt + lf.wl - lf.wg, lf.zl),
             (lf.wt + lf.wl - lf.wg, lf.zg), (lf.wt + lf.wl, lf.zg), (lf.wt + lf.wl, zmax),
             (lf.wt + lf.wts - lf.wbs, zmax), (lf.wt + lf.wts - lf.wbs, lf.zbs), (lf.wt + lf.wts, lf.zbs),
             (lf.wt + lf.wts, lf.zts), (lf.wt, lf.zts), (lf.wt, lf.zt), (0, lf.zt)]
    elif kind == "target":
        v = [(0, zmin), (lf.wt + lf.wl - lf.wbs + lf.wts, zmin), (lf.wt + lf.wl - lf.wbs + lf.wts, lf.zts),
             (lf.wt + lf.wl - lf.wbs, lf.zts), (lf.wt + lf.wl - lf.wbs, lf.zbs), (lf.wt + lf.wl, lf.zbs),
             (lf.wt + lf.wl, lf.zg), (lf.wt + lf.wl - lf.wg, lf.zg), (lf.wt + lf.wl - lf.wg, lf.ztip),
             (lf.wt + lf.wl - lf.wg - lf.wtip, lf.ztip), (lf.wt + lf.wl - lf.wg - lf.wtip, lf.zl),
             (lf.wt, lf.zl), (lf.wt, lf.zt), (0, lf.zt)]
    else:
        v = [(0, lf.zt), (lf.wt, lf.zt), (lf.wt, lf.ztip), (lf.wt + lf.wtip, lf.ztip), (lf.wt + lf.wtip, lf.zl),
             (lf.wt + lf.wl - lf.wg, lf.zl), (lf.wt + lf.wl - lf.wg, lf.zg), (lf.wt + lf.wl, lf.zg),
             (lf.wt + lf.wl, zmax), (lf.wts - lf.wbs, zmax), (lf.wts - lf.wbs, lf.zbs), (lf.wts, lf.zbs),
             (lf.wts, lf.zts), (0, lf.zts)]
    return [(np.float32(x), np.float32(z)) for x, z in v]


def leaf_end_profile(zmin, zthick, rmax, rad_end):
    """The closed leaf pair along the leaf travel, 34 (y, z) vertices: the
    -rmax edge, bank A's rounded end (tip at y = 0 at mid height), bank B's
    end, the +rmax edge."""
    zc = zthick / 2.0
    th = math.asin(min(1.0, zc / rad_end))
    ang = np.linspace(-th, th, 15)
    # the C library's sin and cos (as the C++ reader evaluates them), not
    # numpy's vectorised ones, which may differ in the last binary64 place
    ys = np.array([rad_end * math.sin(a) for a in ang])          # along z, relative to the mid plane
    xs = np.array([rad_end * (1.0 - math.cos(a)) for a in ang])  # retreat of the end from the tip
    out = np.zeros((34, 2), dtype=np.float32)
    out[0] = (-rmax, zmin + zthick)
    out[1] = (-rmax, zmin)
    for i in range(15):                          # bank A, top to bottom, y <= 0
        out[2 + i] = (-xs[i], ys[i] + zc + zmin)
    for i in range(15):                          # bank B, bottom to top, y >= 0
        out[17 + i] = (xs[14 - i], ys[14 - i] + zc + zmin)
    out[32] = (rmax, zmin)
    out[33] = (rmax, zmin + zthick)
    return out


def build_mlc(d: Dynvmlc) -> Mlc:
    """gegs_mlc of a DYNVMLC description (read_egsinp, init.cu:441-611)."""
    m = Mlc()
    f32 = np.float32
    zmin, zthick = f32(d.zmin), f32(d.zthick)
    m.zmin, m.zthick, m.rmax = float(zmin), float(zthick), float(d.rmax)
    shapes = {"full": _vertices("full", d.full, zmin, zthick), "target": _vertices("target", d.target, zmin, zthick),
              "isocenter": _vertices("isocenter", d.isocenter, zmin, zthick)}
    order = ["full"] * d.n_full + ["target", "isocenter"] * (d.n_half // 2) + ["full"] * d.n_full
    if len(order) != NLEAF:
        raise ValueError("DYNVMLC: %d leaves, expected %d" % (len(order), NLEAF))
    width = {"full": f32(d.full.wl), "target": f32(d.target.wl), "isocenter": f32(d.isocenter.wl)}
    X = f32(d.x_start)
    gap = f32(d.air_gap)
    x_mid_1 = x_mid_2 = None
    for i, kind in enumerate(order):
        if i == d.n_full:
            x_mid_1 = X
        if i == d.n_full + d.n_half:
            x_mid_2 = X
        for j, (vx, vz) in enumerate(shapes[kind]):
            m.cross_xz[i * 16 + j].x = float(f32((X + vx) / zmin * vz))
            m.cross_xz[i * 16 + j].y = float(vz)
        X = f32(X + width[kind] + gap)
    x_end = f32(X - gap + f32(d.full.wg))
    m.x_start, m.x_end = float(d.x_start), float(x_end)
    m.x_mid_1, m.x_mid_2 = float(x_mid_1), float(x_mid_2)
    m.air_gap = float(gap)
    m.wl_full, m.wl_target, m.wl_isocenter = float(d.full.wl), float(d.target.wl), float(d.isocenter.wl)
    m.wt = float(d.full.wt)
    # default openings in the 15th slot of every leaf (init.cu:497-515)
    i = 0
    for neg, pos, num in d.openings:
        for _ in range(num):
            m.cross_xz[14 + i * 16].x = neg
            m.cross_xz[14 + i * 16].y = pos
            i += 1
    zt = f32(zmin + zthick)
    bxz = [(f32(d.x_start) / zmin * zt, zt), (f32(d.x_start), zmin), (x_end, zmin), (x_end / zmin * zt, zt)]
    byz = [(-d.rmax, zt), (-d.rmax, zmin), (d.rmax, zmin), (d.rmax, zt)]
    for k in range(4):
        m.boundary_xz[k].x, m.boundary_xz[k].y = float(bxz[k][0]), float(bxz[k][1])
        m.boundary_yz[k].x, m.boundary_yz[k].y = float(byz[k][0]), float(byz[k][1])
    prof = leaf_end_profile(float(zmin), float(zthick), d.rmax, d.rad_end)
    for k in range(34):
        m.cross_yz[k].x, m.cross_yz[k].y = float(prof[k, 0]), float(prof[k, 1])
    return m


def _leaf(wl, wt, wg, wtip, wts, wbs, ztip, zl, zt, zg, zts, zbs):
    return Leaf(wl, wt, wg, wtip, wts, wbs, ztip, zl, zt, zg, zth=zg, zbh=zg, holepos=0.0, zts=zts, zbs=zbs)


# synthetic 120-leaf MLC: leaf pitch 0.48 cm (outer) and 0.24 cm (inner) at
# zmin = 48 cm, i.e. 1.0 cm and 0.5 cm at a 100 cm isocentre; leaves 6.9 cm
# high, a 0.006 cm air gap between neighbours.  The leaves are plain focused
# slabs (no tongue, groove, tip or protruding rail: those widths are 0 and
# the step heights coincide with the leaf faces), so the 14-vertex
# cross-sections degenerate to focused quadrilaterals.  The reference's leaf
# tracking assumes air between neighbouring leaves at every height and a leaf
# lookup from each leaf's nominal span; tongue-and-groove dimensions that
# keep both true need the vendor drawing, which is not available.
MILLENNIUM_LIKE = Dynvmlc(
    rmax=20.0, zmin=48.0, zthick=6.9,
    full=_leaf(0.474, 0.0, 0.0, 0.0, 0.0, 0.0, 48.0, 48.0, 48.0, 54.9, 54.9, 54.9),
    target=_leaf(0.234, 0.0, 0.0, 0.0, 0.0, 0.0, 54.9, 54.9, 48.0, 54.9, 48.0, 54.9),
    isocenter=_leaf(0.234, 0.0, 0.0, 0.0, 0.0, 0.0, 48.0, 48.0, 48.0, 54.9, 54.9, 54.9),
    n_full=10, n_half=40, x_start=0.0,
    air_gap=0.006, rad_end=7.0, openings=[(0.0, 0.0, 60)])


def millennium_like():
    """The synthetic MLC with its leaf rows centred on the beam axis."""
    d = dataclasses.replace(MILLENNIUM_LIKE)
    # centre the leaf rows on the axis: total width at zmin
    total = 2 * d.n_full * (d.full.wl + d.air_gap) + (d.n_half // 2) * (d.target.wl + d.isocenter.wl + 2 * d.air_gap)
    d.x_start = -0.5 * total
    return d


# ------------------------------------------------------------------ egsinp DYNVMLC block
CM_TITLE = "*********** start of CM DYNVMLC"


def write_dynvmlc(path, d: Dynvmlc):
    """The DYNVMLC lines read_egsinp consumes (init.cu:389-465), in order."""
    def leaf_line_a(lf):
        return "%g, %g, %g, %g, %g, %g," % (lf.wl, lf.wt, lf.wg, lf.wtip, lf.wts, lf.wbs)
    lines = [CM_TITLE + " with identifier MLC *********",
             "%g, RMAX" % d.rmax, "Varian-like MLC (synthetic)", "0, ZFOCUS mode",
             "%g, ZMIN" % d.zmin, "%g, ZTHICK" % d.zthick,
             leaf_line_a(d.full),
             "%g, %g, %g, %g, %g, %g," % (d.full.ztip, d.full.zl, d.full.zt, d.full.zg, d.full.zth, d.full.zbh),
             "%g," % d.full.holepos, "%g, %g," % (d.full.zts, d.full.zbs),
             leaf_line_a(d.target),
             "%g, %g, %g, %g," % (d.target.zts, d.target.zbs, d.target.zth, d.target.zbh),
             "%g," % d.target.holepos,
             "%g, %g, %g, %g," % (d.target.zt, d.target.zg, d.target.zl, d.target.ztip),
             leaf_line_a(d.isocenter),
             "%g, %g, %g, %g, %g, %g," % (d.isocenter.ztip, d.isocenter.zl, d.isocenter.zt, d.isocenter.zg,
                                          d.isocenter.zth, d.isocenter.zbh),
             "%g," % d.isocenter.holepos, "%g, %g," % (d.isocenter.zts, d.isocenter.zbs),
             "%d, NUM FULL LEAVES LEFT" % d.n_full, "%d, NUM HALF LEAVES" % d.n_half,
             "%d, NUM FULL LEAVES RIGHT" % d.n_full, "%g, START" % d.x_start, "%g, LEAFGAP" % d.air_gap,
             "0, ENDTYPE", "%g, LEAFRADIUS" % d.rad_end, "0, ZFOCUS of leaf sides"]
    lines += ["%g, %g, %d" % o for o in d.openings]
    lines += ["0.7, 0.01, 0, 0", "0, ", "", d.medium]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _nums(line, n):
    vals = [t for t in line.replace(",", " ").split()]
    out = []
    for t in vals:
        try:
            out.append(float(t))
        except ValueError:
            break
    if len(out) < n:
        raise ValueError("DYNVMLC: expected %d numbers in %r" % (n, line))
    return out[:n]


def read_dynvmlc(path) -> Dynvmlc:
    """Parse the DYNVMLC block the way read_egsinp does (line by line after
    the CM title; comments after the numbers are ignored)."""
    with open(path) as f:
        lines = f.read().splitlines()
    try:
        k = next(i for i, l in enumerate(lines) if l.startswith(CM_TITLE[:31]))
    except StopIteration:
        raise ValueError("no DYNVMLC block in %s" % path)
    it = iter(lines[k + 1:])
    rmax = _nums(next(it), 1)[0]
    next(it)
    next(it)
    zmin = _nums(next(it), 1)[0]
    zthick = _nums(next(it), 1)[0]
    fa = _nums(next(it), 6)
    fb = _nums(next(it), 6)
    fh = _nums(next(it), 1)[0]
    fr = _nums(next(it), 2)
    full = Leaf(*fa, ztip=fb[0], zl=fb[1], zt=fb[2], zg=fb[3], zth=fb[4], zbh=fb[5], holepos=fh, zts=fr[0], zbs=fr[1])
    ta = _nums(next(it), 6)
    tb = _nums(next(it), 4)
    th = _nums(next(it), 1)[0]
    tc = _nums(next(it), 4)
    target = Leaf(*ta, ztip=tc[3], zl=tc[2], zt=tc[0], zg=tc[1], zth=tb[2], zbh=tb[3], holepos=th, zts=tb[0], zbs=tb[1])
    ia = _nums(next(it), 6)
    ib = _nums(next(it), 6)
    ih = _nums(next(it), 1)[0]
    ir = _nums(next(it), 2)
    iso = Leaf(*ia, ztip=ib[0], zl=ib[1], zt=ib[2], zg=ib[3], zth=ib[4], zbh=ib[5], holepos=ih, zts=ir[0], zbs=ir[1])
    n_full = int(_nums(next(it), 1)[0])
    n_half = int(_nums(next(it), 1)[0])
    next(it)
    x_start = _nums(next(it), 1)[0]
    gap = _nums(next(it), 1)[0]
    next(it)
    rad = _nums(next(it), 1)[0]
    next(it)
    openings, n = [], 0
    while n < NLEAF:
        a, b, c = _nums(next(it), 3)
        openings.append((a, b, int(c)))
        n += int(c)
    next(it)
    next(it)
    next(it)
    medium = next(it).strip()
    return Dynvmlc(rmax, zmin, zthick, full, target, iso, n_full, n_half, x_start, gap, rad, openings, medium)


# ------------------------------------------------------------------ egsinp JAWS block (SECJAWS)
SECJAWS_TITLE = "*********** start of CM JAWS with identifier SECJAWS"   # the 52 characters compared


def read_egsinp_secjaws(path):
    """(Jaws, medium) of the JAWS component SECJAWS (read_egsinp_secjaws,
    init.cu:612-689): after the line that starts with the 52-character CM
    title, RMAX on the next line, three lines skipped, "ZMIN, ZMAX" of the
    first (Y) pair, one line skipped, "ZMIN, ZMAX" of the second (X) pair, two
    lines skipped, then the jaw medium's name.  None when the file has no such
    block (the reference then keeps its zero-initialised jaws)."""
    with open(path, errors="replace") as f:
        lines = f.read().splitlines()
    try:
        k = next(i for i, l in enumerate(lines) if l[:52] == SECJAWS_TITLE)
    except StopIteration:
        return None
    rmax = _nums(lines[k + 1], 1)[0]
    zminy, zmaxy = _nums(lines[k + 5], 2)
    zminx, zmaxx = _nums(lines[k + 7], 2)
    medium = lines[k + 10].strip() if k + 10 < len(lines) else ""
    return build_jaws(rmax, zminy, zmaxy, zminx, zmaxx), medium


def write_egsinp_secjaws(path, jaws, medium="W700ICRU", append=False):
    """A BEAMnrc JAWS component with the lines read_egsinp_secjaws takes
    (written after an existing DYNVMLC block when append is set: the
    reference reads both from one .egsinp file)."""
    lines = [SECJAWS_TITLE + "  ***********",
             "%g, RMAX" % jaws.rmax, "SECJAWS (synthetic)", "2, # PAIRED BARS OR JAWS", "Y",
             "%g, %g, 0, 0, 0, 0," % (jaws.zminy, jaws.zmaxy), "X",
             "%g, %g, 0, 0, 0, 0," % (jaws.zminx, jaws.zmaxx),
             "0.7, 0.01, 0, 0,", "0.7, 0.01, 0, 0, 0", medium]
    with open(path, "a" if append else "w") as f:
        f.write("\n".join(lines) + "\n")


# ------------------------------------------------------------------ fields (.mlcinfo)
@dataclasses.dataclass
class MlcField:
    angle: float
    x1: float
    x2: float
    y1: float
    y2: float
    mu: float
    leaf_file: str
    positions: np.ndarray | None = None     # 120 leaf positions (cm at the isocentre)


def write_leaf_file(path, positions):
    """A Varian leaf file: 14 header lines, then 120 lines whose value starts
    at column 11 (read_mlcinp reads after skipping 10 characters)."""
    pos = np.asarray(positions, dtype=np.float64)
    head = ["File Rev = H", "Treatment = Dynamic Dose", "Last Name = synthetic", "First Name = ",
            "Patient ID = 0", "Number of Fields = 1", "Number of Leaves = 120", "Tolerance = 0.50",
            "Field = 1", "Index = 0.0000", "Carriage Group = 1", "Operator = ", "Collimator = 0.0", ""]
    rows = []
    for j in range(120):
        bank = "A" if j < 60 else "B"
        rows.append("Leaf%3d%s=%8.4f" % (j % 60 + 1, bank, pos[j]))
    with open(path, "w") as f:
        f.write("\n".join(head + rows) + "\n")


def read_leaf_file(path):
    with open(path) as f:
        lines = f.read().splitlines()
    vals = [float(lines[14 + j][10:].split()[0].strip("=")) for j in range(120)]
    return np.array(vals, dtype=np.float32)


def write_mlcinfo(path, fields):
    rows = ["MLC fields (synthetic)", "%d" % len(fields)]
    for i, f in enumerate(fields):
        rows.append("%d, %g, %g, %g, %g, %g, %g, %s" % (i + 1, f.angle, f.x1, f.x2, f.y1, f.y2, f.mu, f.leaf_file))
    with open(path, "w") as fo:
        fo.write("\n".join(rows) + "\n")


def read_mlcinfo(path, data_dir=""):
    """Fields of a .mlcinfo list (read_mlcinp, init.cu:1414-1470)."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[1].split(",")[0])
    out = []
    for line in lines[2:2 + n]:
        parts = [t.strip() for t in line.split(",")]
        vals = [float(t) for t in parts[1:7]]
        out.append(MlcField(*vals, leaf_file=data_dir + ",".join(parts[7:]).strip()))
    return out


def random_aperture(seed, half=5.0, smooth=3):
    """120 leaf positions of an irregular aperture inside a 2 half x 2 half
    field: pairs whose projected width lies outside the field stay closed,
    the others open to smoothly varying seeded positions: the +x leaf end at
    +pos[j], the -x end at -pos[j+60], both within `half` (the leaf-file
    order of init.cu:1607-1613)."""
    rng = np.random.default_rng(seed)
    centers = np.concatenate([np.arange(10) * 1.0 - 19.5, np.arange(40) * 0.5 - 9.75, np.arange(10) * 1.0 + 10.5])
    a = rng.uniform(0.3 * half, half, NLEAF)
    b = rng.uniform(0.3 * half, half, NLEAF)
    k = np.ones(smooth) / smooth
    a = np.convolve(a, k, mode="same")
    b = np.convolve(b, k, mode="same")
    pos = np.zeros(2 * NLEAF, dtype=np.float32)
    inside = np.abs(centers) < half
    pos[:NLEAF] = np.where(inside, b, 0.0)            # +x leaf end at +pos[j]
    pos[NLEAF:] = np.where(inside, a, 0.0)            # -x leaf end at -pos[j+60]
    return pos


# ------------------------------------------------------------------ block and wedge
M_WEDGE = 4
M_BLOCK = 5

_PYRAMIDS_TITLE = "*********** start of CM PYRAMIDS with identifier "


class BlockWedge(ctypes.Structure):
    """ctypes mirror of ``gegs_block_wedge`` (include/gegs.h): the block's
    two quadrilaterals, the wedge's cross-section (wedge_nver vertices) and
    its bounding rectangle, head frame, z in cm from the source."""
    _fields_ = [("block_xz", Float2 * 4), ("block_yz", Float2 * 4), ("wedge", Float2 * 6),
                ("wedge_bound", Float2 * 4), ("has_block", ctypes.c_int32), ("has_wedge", ctypes.c_int32),
                ("wedge_nver", ctypes.c_int32), ("pad", ctypes.c_int32)]


def _f2(arr, pts):
    for i, (x, y) in enumerate(pts):
        arr[i].x, arr[i].y = float(np.float32(x)), float(np.float32(y))


def block_polygons(coord):
    """The block's cross-sections from its 10 PYRAMIDS values zmin, zmax,
    XFP, XBP, XFN, XBN, YFP, YBP, YFN, YBN (init.cu:828-844): (XFP, zmin),
    (XBP, zmax), (XBN, zmax), (XFN, zmin) and the same with the Y values."""
    c = [np.float32(v) for v in coord]
    if len(c) != 10:
        raise ValueError("a block needs 10 values")
    xz = [(c[2], c[0]), (c[3], c[1]), (c[5], c[1]), (c[4], c[0])]
    yz = [(c[6], c[0]), (c[7], c[1]), (c[9], c[1]), (c[8], c[0])]
    return np.array(xz, dtype=np.float32), np.array(yz, dtype=np.float32)


def wedge_polygon(rmax, layers):
    """The wedge's cross-section and bounding rectangle from its layers, each
    zmin, zmax, XFP, XBP, XFN, XBN (read_egsinp_wedge, init.cu:752-778).  The
    reference traces the slanted side of the layer stack: when the + side of
    the first layer is vertical (XFP == XBP) the - side's front values, else
    (XFN == XBN) the + side's; in the direction in which the first layer's
    front and back differ, from the first layer up or from the last layer
    down, one vertex per layer, closed by a vertex on the last (first) layer's
    far face and one that pairs the first traced x with the far z.  Returns
    (vertices [layers + 2, 2], rectangle [4, 2])."""
    c = np.asarray(layers, dtype=np.float32).reshape(-1)
    n = len(c) // 6
    if n < 1 or n > 4:
        raise ValueError("The number of layers %d is larger than 4!!" % n)   # error 9005 of the reference
    if c[2] - c[3] == 0:
        start, step, offset = (4, 6, 4) if c[4] - c[5] > 0 else (6 * n - 1, -6, 4)
    elif c[4] - c[5] == 0:
        start, step, offset = (2, 6, 2) if c[2] - c[3] < 0 else (6 * n - 3, -6, 2)
    else:
        raise ValueError("neither side of the wedge's first layer is vertical")
    first_x = c[start]
    pts = [None]
    while 0 <= start < 6 * n:
        pts.append((c[start], c[start - offset]))
        start += step
    start = start - step + step // 6
    pts.append((c[start], c[start - offset]))
    pts[0] = (first_x, c[start - offset])
    bound = [(-rmax, c[0]), (rmax, c[0]), (rmax, c[6 * n - 5]), (-rmax, c[6 * n - 5])]
    return np.array(pts, dtype=np.float32), np.array(bound, dtype=np.float32)


def build_block_wedge(block=None, wedge=None):
    """gegs_block_wedge from block = 10 PYRAMIDS values (or None) and wedge =
    (rmax, layers) (or None)."""
    bw = BlockWedge()
    if block is not None:
        xz, yz = block_polygons(block)
        _f2(bw.block_xz, xz)
        _f2(bw.block_yz, yz)
        bw.has_block = 1
    if wedge is not None:
        geom, bound = wedge_polygon(*wedge)
        _f2(bw.wedge, geom)
        _f2(bw.wedge_bound, bound)
        bw.wedge_nver = len(geom)
        bw.has_wedge = 1
    return bw


def _numbers(line, n):
    """The first n comma-separated numbers of a line (the reference reads
    each with ``>> value >> ch``, the separator going into ch)."""
    toks = [t for t in line.replace("\t", " ").split(",")]
    out = []
    for t in toks:
        t = t.strip()
        if not t:
            continue
        out.append(float(t.split()[0]))
        if len(out) == n:
            return out
    raise ValueError("expected %d numbers in %r" % (n, line))


def _find_pyramids(lines, ident):
    title = (_PYRAMIDS_TITLE + ident)[:54]
    for i, line in enumerate(lines):
        if line[:54] == title:
            return i
    return None


def read_egsinp_block(path):
    """(10 PYRAMIDS values, medium name) of the block of a BEAMnrc input, or
    None when the file has no PYRAMIDS component named BLOCK
    (read_egsinp_block, init.cu:793-905): after the title line, RMAX, two
    lines skipped, the 10 values, two lines skipped, the medium line."""
    with open(path) as f:
        lines = f.read().splitlines()
    i = _find_pyramids(lines, "BLOCK")
    if i is None:
        return None
    coord = _numbers(lines[i + 4], 10)
    return coord, lines[i + 7]


def read_egsinp_wedge(path):
    """(rmax, layers [n, 6], medium name) of the wedge of a BEAMnrc input, or
    None without a PYRAMIDS component named WEDGE (read_egsinp_wedge,
    init.cu:692-791): after the title line, RMAX, a line skipped, the number
    of layers (at most 4), one line of 6 values per layer, two lines
    skipped, the medium line."""
    with open(path) as f:
        lines = f.read().splitlines()
    i = _find_pyramids(lines, "WEDGE")
    if i is None:
        return None
    rmax = _numbers(lines[i + 1], 1)[0]
    n = int(_numbers(lines[i + 3], 1)[0])
    if n > 4:
        raise ValueError("The number of layers %d is larger than 4!!" % n)
    layers = [_numbers(lines[i + 4 + j], 6) for j in range(n)]
    return rmax, np.array(layers, dtype=np.float32), lines[i + 4 + n + 2]


def write_egsinp_pyramids(path, block=None, wedge=None, block_medium="W700ICRU", wedge_medium="W700ICRU"):
    """A BEAMnrc-style input holding a PYRAMIDS component BLOCK (10 values)
    and/or WEDGE ((rmax, layers)), in the line layout the readers above
    expect."""
    out = ["BEAMnrc-style input with PYRAMIDS components"]
    if block is not None:
        out += [_PYRAMIDS_TITLE + "BLOCK", "10.0, RMAX", "block", "1, # LAYERS",
                ", ".join("%.6g" % v for v in block) + ",", "0, 0, 0, 0,", "0, 0, 0, 0,", block_medium]
    if wedge is not None:
        rmax, layers = wedge
        layers = np.asarray(layers, dtype=np.float64).reshape(-1, 6)
        out += [_PYRAMIDS_TITLE + "WEDGE", "%.6g, RMAX" % rmax, "wedge", "%d, # LAYERS" % len(layers)]
        out += [", ".join("%.6g" % v for v in row) + "," for row in layers]
        out += ["0, 0, 0, 0,", "0, 0, 0, 0,", wedge_medium]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def block_wedge_from_run_input(ri):
    """(gegs_block_wedge, block medium, wedge medium) of a run input
    (runinput.RunInput) whose [Block] and/or [Wedge] module is On, or None
    when both are off.  As the reference does (CUDA_EGS.cu:1231-1275,
    init.cu:1327-1329), each file is the data directory followed by the
    file property, and the medium is the line the component names."""
    from .runinput import get
    if not (ri.block or ri.wedge):
        return None
    block = wedge = None
    bmed = wmed = None
    if ri.block:
        path = ri.data_directory + get(ri.sections, "Block", "block file").strip().strip('"')
        r = read_egsinp_block(path)
        if r is None:
            raise ValueError("no PYRAMIDS component BLOCK in %s" % path)
        block, bmed = r
    if ri.wedge:
        path = ri.data_directory + get(ri.sections, "Wedge", "wedge file").strip().strip('"')
        r = read_egsinp_wedge(path)
        if r is None:
            raise ValueError("no PYRAMIDS component WEDGE in %s" % path)
        wedge, wmed = (r[0], r[1]), r[2]
    return build_block_wedge(block, wedge), bmed, wmed
