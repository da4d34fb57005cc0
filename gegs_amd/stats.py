"""Dose statistics: uncertainty estimators, dose efficiency, z-score chi^2 and
3-D gamma index.  These replace the reference's measurement comparisons
(output.c:511-767), which need measured files that do not exist here."""
from __future__ import annotations

import numpy as np


def batch_sigma(dose, dose2, n_batches):
    """Uncertainty of the total from time-sliced batches: dose = sum_b D_b,
    dose2 = sum_b D_b^2.  Returns sigma(total) per voxel."""
    nb = float(n_batches)
    var_b = (dose2 - dose * dose / nb) / (nb - 1.0)      # variance of one batch
    return np.sqrt(np.maximum(var_b, 0.0) * nb)


def history_sigma(dose, dose2, n_hist):
    """History-by-history uncertainty of the total: dose2 = sum_h d_h^2."""
    n = float(n_hist)
    var = (dose2 - dose * dose / n) / (n - 1.0)
    return np.sqrt(np.maximum(var, 0.0) * n)


def mean_rel_sigma(dose, sigma, frac=0.5):
    """Mean relative uncertainty over voxels with dose > frac * max."""
    sel = dose > frac * dose.max()
    return float(np.mean(sigma[sel] / dose[sel])), int(sel.sum())


def efficiency(mean_rel, seconds):
    """Dose efficiency 1 / (sigma^2 T)."""
    return 1.0 / (mean_rel * mean_rel * seconds)


def zscore_chi2(d1, s1, d2, s2, mask):
    """Per-voxel z = (d1 - d2) / sqrt(s1^2 + s2^2) over mask; returns
    (chi2 per degree of freedom, mean z, number of voxels)."""
    den = np.sqrt(s1[mask] ** 2 + s2[mask] ** 2)
    ok = den > 0
    z = (d1[mask][ok] - d2[mask][ok]) / den[ok]
    return float(np.mean(z * z)), float(np.mean(z)), int(z.size)


def gamma_pass_rate(ref, ev, spacing, dd=0.01, dta=0.1, threshold=0.1, search=2):
    """Global 3-D gamma: dose difference dd (fraction of max ref), distance to
    agreement dta (cm), evaluated on ref voxels above threshold*max(ref), with
    a search over +-search voxels (trilinear sub-voxel interpolation of the
    evaluated dose along each axis at half-voxel steps).  ref/ev are (nz, ny, nx)."""
    ref = np.asarray(ref, dtype=np.float64)
    ev = np.asarray(ev, dtype=np.float64)
    dmax = ref.max()
    norm_d = dd * dmax
    dz, dy, dx = spacing
    mask = ref > threshold * dmax
    best = np.full(ref.shape, np.inf)
    steps = np.arange(-2 * search, 2 * search + 1) * 0.5
    pad = search + 1
    evp = np.pad(ev, pad, mode="edge")
    nz, ny, nx = ref.shape
    for sz in steps:
        for sy in steps:
            for sx in steps:
                dist2 = (sz * dz) ** 2 + (sy * dy) ** 2 + (sx * dx) ** 2
                if dist2 > (search * min(dx, dy, dz)) ** 2 + 1e-12:
                    continue
                # trilinear sample of ev at offset (sz, sy, sx)
                fz, fy, fx = np.floor(sz), np.floor(sy), np.floor(sx)
                wz, wy, wx = sz - fz, sy - fy, sx - fx
                val = 0.0
                for iz, cz in ((0, 1 - wz), (1, wz)):
                    if cz == 0:
                        continue
                    for iy, cy in ((0, 1 - wy), (1, wy)):
                        if cy == 0:
                            continue
                        for ix, cx in ((0, 1 - wx), (1, wx)):
                            if cx == 0:
                                continue
                            oz, oy, ox = int(fz) + iz + pad, int(fy) + iy + pad, int(fx) + ix + pad
                            val = val + cz * cy * cx * evp[oz:oz + nz, oy:oy + ny, ox:ox + nx]
                g2 = (val - ref) ** 2 / norm_d ** 2 + dist2 / dta ** 2
                np.minimum(best, g2, out=best)
    passed = best[mask] <= 1.0
    return float(passed.mean()), int(mask.sum())
