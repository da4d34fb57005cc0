"""Voxel phantoms: egsphant I/O, ctcreate-style HU -> media conversion, the
region table, and the seeded synthetic phantoms of the benchmark workloads.

Reference behaviour restated here:

* ``ct_to_medium``   ctcreate.c:824-1140 (``CTToMedium``) with the hard-wired
  4-media ramp of ctcreate.c:904-993; CT numbers below the ramp become
  medium 0 = vacuum (quirk Q10, ctcreate.c:1060-1066).
* ``write_egsphant`` / ``read_egsphant``  ctcreate.c:1204-1340 and
  init.cu:1164-1271 (media names, bounds, ``%1d`` media digits, densities).
* ``region_table``   init.cu:909-1051 (``init_regions``): region 0 vacuum,
  regions 1..5 the head-module media, voxels from region 6 on, x fastest;
  rhof = voxel density / medium density, cuts = the medium's AE/AP.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

VAC_AND_OTH = 6
VACUUM = 0xFFFF
F_RAYLEIGH = 0x0001
F_BOUND_COMPTON = 0x0002
F_ATOMIC_RELAXATION = 0x0010
F_PHOTOELECTRON_ANGLE = 0x0020

# Region flags of the transport, as the reference sets them for every region
# (CUDA_EGS.h:44-51, init.cu:935-1000): Rayleigh scattering, atomic
# relaxation, photo-electron angular distribution and bound Compton
# scattering with Doppler broadening.
DEFAULT_FLAGS = F_RAYLEIGH | F_ATOMIC_RELAXATION | F_PHOTOELECTRON_ANGLE | F_BOUND_COMPTON
# The configuration round-1 results were quoted in (Klein-Nishina Compton, no
# relaxation, no shell selection in photo-absorption).
KN_FLAGS = F_RAYLEIGH | F_PHOTOELECTRON_ANGLE

REGION_DTYPE = np.dtype([("med", "<u2"), ("flags", "<u2"), ("rhof", "<f4"),
                         ("pcut", "<f4"), ("ecut", "<f4")])

# default ctcreate ramp (ctcreate.c:904-969): name, CT upper bound, density lo, hi
DEFAULT_RAMP_LOWER = -1024
DEFAULT_RAMP = [
    ("AIR700ICRU", -974, 0.001, 0.044),
    ("LUNG700ICRU", -724, 0.044, 0.302),
    ("ICRUTISSUE700ICRU", 101, 0.302, 1.101),
    ("ICRPBONE700ICRU", 1976, 1.101, 2.088),
]


class Phantom:
    """Rectilinear voxel phantom: bounds (cm), media index per voxel (1-based,
    0 = vacuum) and density per voxel (g/cm^3), x fastest."""

    def __init__(self, media_names, xb, yb, zb, media_idx, density):
        self.media_names = list(media_names)
        self.xb = np.asarray(xb, dtype=np.float32)
        self.yb = np.asarray(yb, dtype=np.float32)
        self.zb = np.asarray(zb, dtype=np.float32)
        self.nx, self.ny, self.nz = len(self.xb) - 1, len(self.yb) - 1, len(self.zb) - 1
        n = self.nx * self.ny * self.nz
        self.media_idx = np.asarray(media_idx, dtype=np.int32).reshape(n)
        self.density = np.asarray(density, dtype=np.float32).reshape(n)
        if self.media_idx.min() < 0 or self.media_idx.max() > len(self.media_names):
            raise ValueError("ERROR (8010): media index outside the media list")

    @property
    def nvox(self):
        return self.nx * self.ny * self.nz

    def padded_bounds(self):
        """Bounds with one extrapolated entry appended, so that the neighbouring-
        voxel read of hownear (quirk Q1, geom.cu:1487-1490) stays defined."""
        out = []
        for b in (self.xb, self.yb, self.zb):
            ext = np.float32(b[-1] + (b[-1] - b[-2]))
            out.append(np.ascontiguousarray(np.append(b, ext), dtype=np.float32))
        return out

    def voxel_mass(self):
        """Voxel masses (g) as in init.cu:1256-1271, in float: the volume
        (dx * dy) * dz, its absolute value, times the density; a mass of 0
        becomes 1."""
        dx = np.diff(self.xb.astype(np.float32))
        dy = np.diff(self.yb.astype(np.float32))
        dz = np.diff(self.zb.astype(np.float32))
        vol = np.abs((dx[None, None, :] * dy[None, :, None]) * dz[:, None, None]).reshape(-1)
        m = (vol * np.asarray(self.density, dtype=np.float32).reshape(-1)).astype(np.float32)
        m[m == 0] = 1.0
        return m


def region_table(phantom, tables, other_media=None, flags=DEFAULT_FLAGS):
    """init_regions (init.cu:909-1051).  ``tables`` is a hatch.MediaTables whose
    ordered names cover the phantom media; ``other_media`` gives the media of
    regions 1..5 (default air, jaws, MLC, block, wedge), by name."""
    c = tables.c
    n = phantom.nvox + VAC_AND_OTH
    reg = np.zeros(n, dtype=REGION_DTYPE)
    reg[0] = (VACUUM, 0, 0.0, 0.0, 0.0)
    other_media = other_media or [phantom.media_names[0]] * 5
    for i in range(1, VAC_AND_OTH):
        m = tables.index(other_media[i - 1])
        reg[i] = (m, flags, 1.0, c.ap[m], c.ae[m])
    # map phantom media (1-based) to table indices
    lut_med = np.full(len(phantom.media_names) + 1, VACUUM, dtype=np.uint16)
    lut_rho = np.zeros(len(phantom.media_names) + 1, dtype=np.float32)
    lut_ap = np.zeros(len(phantom.media_names) + 1, dtype=np.float32)
    lut_ae = np.zeros(len(phantom.media_names) + 1, dtype=np.float32)
    for k, name in enumerate(phantom.media_names):
        m = tables.index(name)
        lut_med[k + 1] = m
        lut_rho[k + 1] = tables.media[m].rho
        lut_ap[k + 1] = c.ap[m]
        lut_ae[k + 1] = c.ae[m]
    mi = phantom.media_idx
    vox = reg[VAC_AND_OTH:]
    vox["med"] = lut_med[mi]
    vac = mi == 0
    vox["flags"] = np.where(vac, 0, flags)
    dens = phantom.density
    with np.errstate(divide="ignore", invalid="ignore"):
        rhof = np.where(dens == 0.0, np.float32(1.0), dens / lut_rho[mi].astype(np.float32))
    vox["rhof"] = np.where(vac, 0.0, rhof).astype(np.float32)
    vox["pcut"] = np.where(vac, 0.0, lut_ap[mi])
    vox["ecut"] = np.where(vac, 0.0, lut_ae[mi])
    return reg


# ------------------------------------------------------------------ ctcreate
def ct_to_medium(hu, ramp=DEFAULT_RAMP, lower=DEFAULT_RAMP_LOWER, fix_q10=False):
    """CTToMedium (ctcreate.c:824-1140): HU -> (1-based medium, density).
    CT values below the ramp become medium 0, vacuum, as in the reference
    (quirk Q10, ctcreate.c:1060-1066); with fix_q10 they take the first ramp
    medium at its lowest density, as values above the ramp take the last.
    The reference holds the ramp's densities as binary32 (ctcreate.c:862-863)
    and computes dlo + (CT - lo) * (dhi - dlo) / (up - lo) in binary32 with
    the density difference taken between the binary32 values
    (ctcreate.c:1074-1095); for the built-in ramp's 0.302..1.101 and
    1.101..2.088 that difference is one unit in the last place away from the
    binary32 rounding of the decimal difference."""
    hu = np.asarray(hu, dtype=np.int32)
    med = np.zeros(hu.shape, dtype=np.int32)
    dens = np.zeros(hu.shape, dtype=np.float32)
    upper = [r[1] for r in ramp]
    below = hu < lower
    above = hu > upper[-1]
    med[above] = len(ramp)
    dens[above] = np.float32(ramp[-1][3])
    if fix_q10:
        med[below] = 1
        dens[below] = np.float32(ramp[0][2])
    rest = ~(below | above)
    assigned = np.zeros(hu.shape, dtype=bool)
    for i, (_, up, dlo, dhi) in enumerate(ramp):
        sel = rest & ~assigned & (hu <= up)
        lo_ct = lower if i == 0 else upper[i - 1]
        span = np.float32(np.float32(dhi) - np.float32(dlo))
        d = np.float32(dlo) + (hu[sel] - lo_ct).astype(np.float32) * span / np.float32(up - lo_ct)
        dens[sel] = d.astype(np.float32)
        med[sel] = i + 1
        assigned |= sel
    return med, dens


# ------------------------------------------------------------------ egsphant
def write_egsphant(path, ph):
    with open(path, "w") as f:
        f.write("%2d\n" % len(ph.media_names))
        for n in ph.media_names:
            f.write(n + "\n")
        f.write(" ".join(["0.25"] * len(ph.media_names)) + "\n")
        f.write("%5d%5d%5d\n" % (ph.nx, ph.ny, ph.nz))
        for b in (ph.xb, ph.yb, ph.zb):
            f.write(" ".join("%.6f" % v for v in b) + "\n")
        mi = ph.media_idx.reshape(ph.nz, ph.ny, ph.nx)
        for k in range(ph.nz):
            for j in range(ph.ny):
                f.write("".join(str(int(v)) for v in mi[k, j]) + "\n")
            f.write("\n")
        de = ph.density.reshape(ph.nz, ph.ny, ph.nx)
        for k in range(ph.nz):
            for j in range(ph.ny):
                f.write(" ".join("%.6f" % v for v in de[k, j]) + "\n")
            f.write("\n")


def read_egsphant(path):
    """egsphant reader following init.cu:1164-1251; errors use the reference's
    8xxx codes."""
    with open(path) as f:
        lines = f.read().split("\n")
    p = 0
    nmed = int(lines[p].split()[0]); p += 1
    names = []
    for _ in range(nmed):
        names.append(lines[p].strip()); p += 1
    p += 1     # estepe line
    try:
        nx, ny, nz = map(int, lines[p].split()[:3]); p += 1
    except ValueError:
        raise ValueError("ERROR (8006): could not read the voxel numbers")
    tok = []

    def take(n, code):
        nonlocal p
        while len(tok) < n:
            if p >= len(lines):
                raise ValueError("ERROR (%d): phantom file truncated" % code)
            tok.extend(lines[p].split()); p += 1
        v = tok[:n]
        del tok[:n]
        return np.array(v, dtype=np.float32)
    xb = take(nx + 1, 8007)
    yb = take(ny + 1, 8008)
    zb = take(nz + 1, 8009)
    mi = np.zeros(nx * ny * nz, dtype=np.int32)
    idx = 0
    for k in range(nz):
        for j in range(ny):
            row = lines[p].strip(); p += 1
            if len(row) < nx:
                raise ValueError("ERROR (8010): could not read media index (y=%d, z=%d)" % (j, k))
            mi[idx:idx + nx] = [int(ch) for ch in row[:nx]]
            idx += nx
        if lines[p].strip() != "":
            raise ValueError("ERROR (8011): expected empty line (z=%d)" % k)
        p += 1
    de = take(nx * ny * nz, 8012)
    return Phantom(names, xb, yb, zb, mi, de)


# ------------------------------------------------------------------ workloads
def water_phantom(n=60, size=30.0, medium="H2O700ICRU"):
    """Homogeneous water cube, n^3 voxels; x,y centred on 0, z from 0."""
    half = size / 2.0
    xb = np.linspace(-half, half, n + 1, dtype=np.float64).astype(np.float32)
    zb = np.linspace(0.0, size, n + 1, dtype=np.float64).astype(np.float32)
    rho = np.float32(1.0)
    return Phantom([medium], xb, xb.copy(), zb, np.ones(n ** 3, dtype=np.int32),
                   np.full(n ** 3, rho, dtype=np.float32))


def synthetic_ct_hu(nx, ny, nz, dx, seed):
    """Seeded synthetic CT (HU): elliptical water body (0 +- 20), two lung
    ellipsoids (-800 +- 50), a spine-like bone cylinder (+900 +- 80) and air
    (-1000) outside.  The body's anterior-posterior axis is y."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5 - nx / 2) * dx
    y = (np.arange(ny) + 0.5 - ny / 2) * dx
    z = (np.arange(nz) + 0.5 - nz / 2) * dx
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ax = 0.42 * nx * dx      # semi-axes of the body ellipse
    ay = 0.30 * ny * dx
    hu = np.full(X.shape, -1000.0)
    body = (X / ax) ** 2 + (Y / ay) ** 2 <= 1.0
    hu[body] = rng.normal(0.0, 20.0, body.sum())
    for sx in (-1.0, 1.0):
        cx = sx * 0.45 * ax
        lung = ((X - cx) / (0.32 * ax)) ** 2 + ((Y + 0.1 * ay) / (0.6 * ay)) ** 2 + (Z / (0.38 * nz * dx)) ** 2 <= 1.0
        hu[lung & body] = rng.normal(-800.0, 50.0, (lung & body).sum())
    spine = (X / (0.12 * ax)) ** 2 + ((Y - 0.62 * ay) / (0.12 * ax)) ** 2 <= 1.0
    hu[spine & body] = rng.normal(900.0, 80.0, (spine & body).sum())
    return np.rint(hu).astype(np.int32)


def ct_phantom(nx=128, ny=128, nz=100, dx=0.25, seed=2):
    """CT phantom of the heterogeneous workloads, converted with the default
    ctcreate ramp; bounds centred on the origin."""
    hu = synthetic_ct_hu(nx, ny, nz, dx, seed)
    med, dens = ct_to_medium(hu.reshape(-1))
    xb = ((np.arange(nx + 1) - nx / 2) * dx).astype(np.float32)
    yb = ((np.arange(ny + 1) - ny / 2) * dx).astype(np.float32)
    zb = ((np.arange(nz + 1) - nz / 2) * dx).astype(np.float32)
    return Phantom([r[0] for r in DEFAULT_RAMP], xb, yb, zb, med, dens)


def dose_to_gy_bin(edep, phantom, path=None, dose_unit=1.6e-8):
    """Dose output of write_dose (output.c:873-903): FP32 x-fastest
    E/(m) * dose_unit, file name <prefix>_DoseResults_NNNxNNNxNNN.bin."""
    m = phantom.voxel_mass()
    out = (np.asarray(edep, dtype=np.float64).astype(np.float32) / m * np.float32(dose_unit)).astype(np.float32)
    if path is not None:
        fname = "%s_DoseResults_%03dx%03dx%03d.bin" % (path, phantom.nx, phantom.ny, phantom.nz)
        out.tofile(fname)
        return out, fname
    return out, None
