"""Dose post-processing of a plan: per-field output factors and MU weighting,
interpolation of the dose onto a CT grid, absolute dose at a point.

Restates the reference's host-side steps in binary32 where the reference
computes in float:

* ``field_output_factor`` -- init.cu:1444-1455: the equivalent square
  sqrt((x2 - x1)(y2 - y1)) of the jaw opening, linearly interpolated in the
  (field size, output factor) table between the first pair of sizes that
  brackets it (lower size exclusive, upper inclusive); None when no pair does
  (the reference then leaves the field's factor unset).
* ``plan_dose`` -- CUDA_EGS.cu:1715-1735: the plan's energy deposit is the sum
  over fields of each field's deposit times its output factor (when output
  factors are on); each field runs MU x histories-per-MU histories
  (CUDA_EGS.cu:377).
* ``interp_dose`` -- output.c:221-315: separable linear interpolation along z,
  then y, then x, onto points start + size * i; the dose of voxel k is placed
  at its lower boundary, points at or outside the first or the last-but-one
  boundary get 0.  (The reference keeps the three passes in one array with
  offsets that keep each pass's inputs intact for grids at least as fine as
  the phantom; separate arrays give the same values.)
* ``abs_dose`` -- output.c:780-872: trilinear interpolation at a point with
  the same boundary placement, the voxel searched from index 1 and clamped to
  [0, n - 2], and the dose of the voxel nearest the point, whose z midpoint
  the reference takes as (z_bounds[k] + x_bounds[k + 1]) / 2 (quirk Q9, kept).
"""
from __future__ import annotations

import os

import numpy as np

f32 = np.float32


def read_output_factor_table(path):
    """Output-factor file (init.cu:1389-1412): a header line, the number of
    entries, then one "field_size output_factor" pair per line."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[1].split()[0])
    sizes, factors = [], []
    for line in lines[2:2 + n]:
        a, b = line.split()[:2]
        sizes.append(float(a))
        factors.append(float(b))
    return np.array(sizes, dtype=np.float32), np.array(factors, dtype=np.float32)


def field_output_factor(x1, x2, y1, y2, sizes, factors):
    """Output factor of a field with jaw opening [x1, x2] x [y1, y2] (cm)."""
    a = f32(np.sqrt(f32(f32(f32(x2) - f32(x1)) * f32(f32(y2) - f32(y1)))))
    for j in range(len(sizes) - 1):
        if a > sizes[j] and a <= sizes[j + 1]:
            return f32(factors[j] + f32(f32(f32(a - sizes[j]) * f32(factors[j + 1] - factors[j]))
                                        / f32(sizes[j + 1] - sizes[j])))
    return None


def plan_dose(field_deposits, output_factors=None):
    """Sum of the fields' energy deposits (FP64), each times its output factor
    when output factors are given."""
    total = np.zeros_like(np.asarray(field_deposits[0], dtype=np.float64))
    for i, d in enumerate(field_deposits):
        d = np.asarray(d, dtype=np.float64)
        total += d * float(output_factors[i]) if output_factors is not None else d
    return total


def histories_for_field(mu, histories_per_mu):
    """Histories of a field: MU x histories per MU, truncated (CUDA_EGS.cu:377)."""
    return int(mu * histories_per_mu)


def _axis_interp(vals, bounds, coords):
    """Linear interpolation of vals[..., k] placed at bounds[k] onto coords
    along the last axis (one pass of interp_dose)."""
    n = vals.shape[-1]
    out = np.zeros(vals.shape[:-1] + (len(coords),), dtype=np.float32)
    for c, t in enumerate(coords):
        t = f32(t)
        if t <= bounds[0] or t >= bounds[n - 1]:
            continue
        for k in range(n - 1):
            if t > bounds[k] and t <= bounds[k + 1]:
                lo, hi = vals[..., k], vals[..., k + 1]
                out[..., c] = lo + f32(t - bounds[k]) * (hi - lo) / f32(bounds[k + 1] - bounds[k])
                break
    return out


def interp_dose(dose, xb, yb, zb, start, sizes, dimen):
    """dose[k, j, i] (float32, nz x ny x nx) on the phantom with bounds xb, yb,
    zb, interpolated onto the points start + sizes * index (dimen points per
    axis); returns a float32 array [dimen_z, dimen_y, dimen_x]."""
    d = np.asarray(dose, dtype=np.float32)
    xb, yb, zb = (np.asarray(b, dtype=np.float32) for b in (xb, yb, zb))
    pts = [np.array([f32(f32(start[a]) + f32(f32(sizes[a]) * f32(i))) for i in range(dimen[a])], dtype=np.float32)
           for a in range(3)]
    # z pass: values along z for every (y, x) column
    vz = _axis_interp(np.moveaxis(d, 0, -1), zb, pts[2])          # [ny, nx, dz]
    vy = _axis_interp(np.moveaxis(vz, 0, -1), yb, pts[1])         # [nx, dz, dy]
    vx = _axis_interp(np.moveaxis(vy, 0, -1), xb, pts[0])         # [dz, dy, dx]
    return vx


def write_interp_dose(path_prefix, grid):
    """The interpolated dose file of output.c:222-312: int32 dimensions x, y,
    z then the float32 values, x fastest; name <prefix>_Interpolate_XXXxYYYxZZZ.bin."""
    nz, ny, nx = grid.shape
    fname = "%s_Interpolate_%03dx%03dx%03d.bin" % (path_prefix, nx, ny, nz)
    with open(fname, "wb") as f:
        np.array([nx, ny, nz], dtype=np.int32).tofile(f)
        np.ascontiguousarray(grid, dtype=np.float32).tofile(f)
    return fname


def _cell(b, n, p):
    i = 1
    while i < n - 2:
        if p > b[i] and p <= b[i + 1]:
            break
        i += 1
    if i == n - 2:
        i = 0 if p <= b[1] else n - 2
    return i


def abs_dose(dose, xb, yb, zb, pos, fix_q9=False):
    """(dose at pos, (i, j, k) of the voxel taken as nearest, its dose).
    fix_q9 takes the z midpoint from the z bounds, (zb[k] + zb[k + 1]) / 2."""
    d = np.asarray(dose, dtype=np.float32)
    nz, ny, nx = d.shape
    xb, yb, zb = (np.asarray(b, dtype=np.float32) for b in (xb, yb, zb))
    px, py, pz = (f32(v) for v in pos)
    i, j, k = _cell(xb, nx, px), _cell(yb, ny, py), _cell(zb, nz, pz)

    def zlin(ii, jj):
        a, b = d[k, jj, ii], d[k + 1, jj, ii]
        return f32(a + f32(f32(pz - zb[k]) * f32(b - a)) / f32(zb[k + 1] - zb[k]))

    a0, a1, a2, a3 = zlin(i, j), zlin(i + 1, j), zlin(i, j + 1), zlin(i + 1, j + 1)
    a0 = f32(a0 + f32(f32(py - yb[j]) * f32(a2 - a0)) / f32(yb[j + 1] - yb[j]))
    a1 = f32(a1 + f32(f32(py - yb[j]) * f32(a3 - a1)) / f32(yb[j + 1] - yb[j]))
    a0 = f32(a0 + f32(f32(px - xb[i]) * f32(a1 - a0)) / f32(xb[i + 1] - xb[i]))
    if px > f32(f32(xb[i] + xb[i + 1]) / f32(2)):
        i += 1
    if py > f32(f32(yb[j] + yb[j + 1]) / f32(2)):
        j += 1
    if not fix_q9 and k + 1 >= len(xb):
        # the reference reads x_bounds[k + 1] here (output.c:835), past the end of
        # the x bounds when the point's z cell is beyond nx - 1: undefined there
        raise ValueError("abs_dose: z cell %d needs x bound %d of %d (reference quirk Q9 reads x_bounds "
                         "for the z midpoint); pass fix_q9=True" % (k, k + 1, len(xb)))
    zhi = zb[k + 1] if fix_q9 else xb[k + 1]           # the reference's z midpoint uses x_bounds (Q9)
    if pz > f32(f32(zb[k] + zhi) / f32(2)):
        k += 1
    return float(a0), (i, j, k), float(d[k, j, i])


def gamma_pass_rate(ref, evl, spacing, dose_crit=0.01, dta=0.1, cutoff=0.1):
    """Global gamma index of the evaluated dose against the reference dose on
    the same voxel grid (arrays [z, y, x], spacing (dx, dy, dz) in cm): for each
    reference voxel with dose above cutoff x max(ref), gamma = min over
    evaluated voxels within 3 x dta of sqrt((dD / (dose_crit x max(ref)))^2 +
    (r / dta)^2); returns the fraction of those voxels with gamma <= 1."""
    ref = np.asarray(ref, dtype=np.float64)
    evl = np.asarray(evl, dtype=np.float64)
    dmax = ref.max()
    norm = dose_crit * dmax
    dx, dy, dz = spacing
    reach = [int(np.ceil(3.0 * dta / h)) for h in (dz, dy, dx)]
    best = np.full(ref.shape, np.inf)
    pad = [(r, r) for r in reach]
    ev = np.pad(evl, pad, mode="constant", constant_values=np.nan)
    nz, ny, nx = ref.shape
    for kz in range(-reach[0], reach[0] + 1):
        for ky in range(-reach[1], reach[1] + 1):
            for kx in range(-reach[2], reach[2] + 1):
                r2 = (kz * dz) ** 2 + (ky * dy) ** 2 + (kx * dx) ** 2
                if r2 > (3.0 * dta) ** 2:
                    continue
                sh = ev[reach[0] + kz:reach[0] + kz + nz, reach[1] + ky:reach[1] + ky + ny,
                        reach[2] + kx:reach[2] + kx + nx]
                g2 = ((sh - ref) / norm) ** 2 + r2 / dta ** 2
                best = np.fmin(best, g2)
    sel = ref > cutoff * dmax
    return float(np.mean(best[sel] <= 1.0))


# ------------------------------------------------------------------ what write_dose reads
def host_bounds(b):
    """The voxel coordinates the reference's output steps use: after the
    bounds go to the device, entry k < n of the host array becomes the
    centre (b[k] + b[k + 1]) / 2 of voxel k, in float; entry n keeps the last
    boundary (init.cu:1289-1294)."""
    b = np.asarray(b, dtype=np.float32).copy()
    n = len(b) - 1
    b[:n] = (b[:n] + b[1:]) / f32(2)
    return b


def voxel_dose(edep, mass, dose_unit):
    """write_dose's values (output.c:888-891): (float) E / m * dose_unit in
    float, per voxel."""
    e = np.asarray(edep, dtype=np.float64).astype(np.float32)
    return ((e / np.asarray(mass, dtype=np.float32)) * f32(dose_unit)).astype(np.float32)


def dose_file_name(prefix, nx, ny, nz):
    return "%s_DoseResults_%03dx%03dx%03d.bin" % (prefix, nx, ny, nz)


def abs_dose_report(dose, xb, yb, zb, pos, calibrated, measured, fix_q9=False):
    """The lines output_AbsDos logs (output.c:780-872) for the point pos of
    the dose (float32, [z, y, x]) on the host coordinates xb, yb, zb."""
    val, (i, j, k), vox = abs_dose(dose, xb, yb, zb, pos, fix_q9=fix_q9)
    val, vox, meas = f32(val), f32(vox), f32(measured)
    px, py, pz = (float(f32(v)) for v in pos)
    if calibrated:
        if meas == 0:
            return ["Absolute dose at (%.2f, %.2f, %.2f) = %.1f cGy" % (px, py, pz, val),
                    "   voxel dose of (i=%2d, j=%2d, k=%2d) = %.1f cGy" % (i, j, k, vox)]
        d0 = f32(abs(f32(val - meas)) / meas)
        d1 = f32(abs(f32(vox - meas)) / meas)
        return ["Absolute dose at (%.2f, %.2f, %.2f):" % (px, py, pz),
                "  measured = %.1f cGy, calculated = %.1f cGy, difference = %.1f %%" % (meas, val, f32(d0 * f32(100))),
                "  voxel dose of index (%2d, %2d, %2d) = %.1f cGy, difference = %.1f %%" % (i, j, k, vox,
                                                                                              f32(d1 * f32(100)))]
    return ["Absolute dose at (%.2f, %.2f, %.2f) = %.3E cGy" % (px, py, pz, val),
            "   voxel dose of (i=%2d, j=%2d, k=%2d) = %.3E cGy" % (i, j, k, vox)]


def reldose_points_nearest(dose, xb, yb, zb, start, sizes, dimen):
    """output_RelDos_II's sampling (output.c:603-624): point (i, j, k) at
    start + sizes * index takes the dose of voxel int((p - b[0]) / (b[1] -
    b[0]) + 0.1) per axis, in float, on the host coordinates (voxel
    centres).  A point whose voxel lies outside the phantom is an error (the
    reference reads outside its dose array there)."""
    d = np.asarray(dose, dtype=np.float32)
    nz, ny, nx = d.shape
    b = [np.asarray(v, dtype=np.float32) for v in (xb, yb, zb)]
    n = (nx, ny, nz)
    out = np.zeros(int(dimen[0]) * int(dimen[1]) * int(dimen[2]), dtype=np.float32)
    for k in range(dimen[2]):
        for j in range(dimen[1]):
            for i in range(dimen[0]):
                idx = []
                for a, t in enumerate((i, j, k)):
                    p = f32(f32(start[a]) + f32(f32(sizes[a]) * f32(t)))
                    q = f32(f32(p - b[a][0]) / f32(b[a][1] - b[a][0]))
                    v = int(float(q) + 0.1)            # float + double, truncated toward zero
                    if not 0 <= v < n[a]:
                        raise ValueError("relative dose point (%d, %d, %d) lies outside the phantom on axis %d"
                                         % (i, j, k, a))
                    idx.append(v)
                out[i + dimen[0] * j + dimen[0] * dimen[1] * k] = d[idx[2], idx[1], idx[0]]
    return out


def reldose_points_interp(dose, xb, yb, zb, start, sizes, dimen):
    """output_RelDos's sampling (output.c:317-424): separable linear
    interpolation along z, y, x of the voxel doses placed at the host
    coordinates; None when a point lies at or outside the first or the
    last-but-one coordinate (the reference then writes no file)."""
    g = np.asarray(dose, dtype=np.float32)
    b = [np.asarray(v, dtype=np.float32) for v in (xb, yb, zb)]
    nz, ny, nx = g.shape
    n = (nx, ny, nz)
    for a in range(3):
        end = f32(f32(start[a]) + f32(f32(sizes[a]) * f32(dimen[a] - 1)))
        if f32(start[a]) <= b[a][0] or end >= b[a][n[a] - 1]:
            return None
    grid = interp_dose(g, b[0], b[1], b[2], start, sizes, dimen)
    return np.ascontiguousarray(grid, dtype=np.float32).reshape(-1)


def _read_measured(path, n):
    vals = []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if t:
                vals.append(float(np.float32(float(t[0]))))
            if len(vals) == n:
                break
    if len(vals) < n:
        raise ValueError('the measured data file "%s" holds %d of the %d values' % (path, len(vals), n))
    return np.array(vals, dtype=np.float32)


def compare_measured(calc, meas, dimen, start, sizes, isocenter, norm_type, array_type, gamma_len, gamma_dose):
    """The comparison both relative-dose outputs make (output.c:444-593):
    normalisation (1: each to its maximum; 2: each to its value at the
    isocentre's point; else the measured values x 100), then every point
    whose measured value exceeds 10% of the measured maximum is compared:
    PTW729 by the dose difference alone, otherwise by the dose difference or,
    failing that, a gamma-like test against the calculated values of the
    neighbouring points along x and z (distance 0.25 cm).  Returns (rows of
    (measured, calculated, pass), total, passed, percent)."""
    calc = np.asarray(calc, dtype=np.float32).copy()
    meas = np.asarray(meas, dtype=np.float32).copy()
    n = len(calc)
    mmax = f32(0)
    if norm_type == 1:
        cmax = f32(0)
        for i in range(n):
            mmax = max(mmax, meas[i])
            cmax = max(cmax, calc[i])
        meas = (meas / mmax).astype(np.float32)
        calc = (calc / cmax).astype(np.float32)
        mmax = f32(1)
    elif norm_type == 2:
        for i in range(n):
            mmax = max(mmax, meas[i])
        center = 0
        if f32(sizes[0]) != 0:
            center += int(float(f32(f32(f32(isocenter[0]) - f32(start[0])) / f32(sizes[0]))) + 0.1)
        if f32(sizes[1]) != 0:
            center += int(float(f32(f32(f32(isocenter[1]) - f32(start[1])) / f32(sizes[1]))) + 0.1) * dimen[0]
        if f32(sizes[2]) != 0:
            center += int(float(f32(f32(f32(isocenter[2]) - f32(start[2])) / f32(sizes[2]))) + 0.1) * dimen[0] * dimen[1]
        if not 0 <= center < n:
            raise ValueError("the isocentre's point %d lies outside the %d relative-dose points" % (center, n))
        mc, cc = meas[center], calc[center]
        meas = (meas / mc).astype(np.float32)
        calc = (calc / cc).astype(np.float32)
        mmax = f32(mmax / mc)
    else:
        meas = (meas * f32(100)).astype(np.float32)
        for i in range(n):
            mmax = max(mmax, meas[i])
    thr = f32(mmax * f32(0.1))
    glen, gdos = f32(gamma_len), f32(gamma_dose)
    rows, total, count = [], 0, 0

    def near(c, m):
        v = f32(f32(abs(f32(c - m)) / m) / gdos)
        v = f32(v * v)
        return float(np.sqrt(float(v) + 0.25 * 0.25 / float(glen) / float(glen))) <= 1.0

    for i in range(n):
        p = 0
        m = meas[i]
        if m > thr:
            total += 1
            if array_type == 1:
                if f32(abs(f32(calc[i] - m)) / m) <= gdos:
                    p = 1
                    count += 1
                else:
                    p = -1
            else:
                if f32(abs(f32(calc[i] - m)) / m) <= gdos:
                    p += 1
                iz = i // dimen[0]
                ix = i - iz * dimen[0]
                if ix > 0 and p <= 0 and near(calc[i - 1], m):
                    p += 1
                if ix + 1 < dimen[0] and p <= 0 and near(calc[i + 1], m):
                    p += 1
                if iz > 0 and near(calc[i - dimen[0]], m):         # no p <= 0 test here (output.c:570)
                    p += 1
                if iz + 1 < dimen[2] and p <= 0 and near(calc[i + dimen[0]], m):
                    p += 1
                if p > 0:
                    count += 1
                else:
                    p = -1
        rows.append((float(m), float(calc[i]), p))
    percent = f32(f32(count) / f32(total)) if total else f32(np.nan)
    return rows, total, count, percent


def write_relative_dose(prefix, calc, dimen, measured_path=None, **compare):
    """<prefix>_RelativeDose_XXXxYYYxZZZ.txt: "measured\\t calculated\\t pass"
    rows and the pass-rate line, or the calculated values alone when the
    measured file cannot be opened (output.c:430-436, :585-592).  Returns
    (file name, log lines)."""
    fname = "%s_RelativeDose_%03dx%03dx%03d.txt" % (prefix, dimen[0], dimen[1], dimen[2])
    log = [fname]
    if measured_path is None or not os.path.exists(measured_path):
        with open(fname, "w") as f:
            for v in np.asarray(calc, dtype=np.float32):
                f.write("%f\n" % v)
        return fname, log + ["Error: dose data file reading error!"]
    meas = _read_measured(measured_path, len(calc))
    rows, total, count, percent = compare_measured(calc, meas, dimen, **compare)
    summary = "Totally %d points compared, %d passed, percentage = %.0f %%" % (total, count, f32(percent * f32(100)))
    with open(fname, "w") as f:
        for m, c, p in rows:
            f.write("%f\t %f\t %d\n" % (m, c, p))
        f.write(summary + "\n")
    return fname, log + [summary]
