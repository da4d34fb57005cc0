"""Source particles: egsphsp MODE0 phase-space I/O and the seeded synthetic
sources of the benchmark workloads.

* MODE0 layout (read_egsphsp.h:8-28, read_egsphsp.cpp:169-232): a 28-byte
  header (``MODE_RW[5]`` padded to 8, NPPHSP, NPHOTPHSP, EKMAXPHSP, EKMINPHSP,
  NINCPHSP) followed by 28-byte records LATCH, E, X, Y, U, V, WT.
* decoding (CUDA_EGS.cu:56-67): LATCH bit 30 -> electron, bit 29 ->
  positron, else photon; W = +sqrt(1-U^2-V^2) with the sign of WT; |E| is the
  total energy; z is the phase-space plane ``z_phsp``.
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

SOURCE_DTYPE = np.dtype([("e", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                         ("u", "<f4"), ("v", "<f4"), ("w", "<f4"), ("wt", "<f4"),
                         ("charge", "<i4")])
PHSP_DTYPE = np.dtype([("latch", "<u4"), ("e", "<f4"), ("x", "<f4"), ("y", "<f4"),
                       ("u", "<f4"), ("v", "<f4"), ("wt", "<f4")])
RM = 0.5110034


def write_egsphsp(path, rec, nincphsp=None):
    """Write MODE0 records (structured array of PHSP_DTYPE)."""
    rec = np.ascontiguousarray(rec, dtype=PHSP_DTYPE)
    q = ((rec["latch"] >> 30) & 1) | ((rec["latch"] >> 29) & 1)
    nphot = int(np.sum(q == 0))
    ek = np.abs(rec["e"]) - np.where(q != 0, RM, 0.0)
    hdr = struct.pack("<5s3xIIffI", b"MODE0", len(rec), nphot, float(ek.max()) if len(rec) else 0.0,
                      float(ek.min()) if len(rec) else 0.0, nincphsp or len(rec))
    assert len(hdr) == 28
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(rec.tobytes())


def read_egsphsp(path, start=0, count=None):
    """Read MODE0 records [start, start+count); returns (header dict, records)."""
    with open(path, "rb") as f:
        hdr = f.read(28)
        mode, npp, nphot, ekmax, ekmin, ninc = struct.unpack("<5s3xIIffI", hdr)
        if mode != b"MODE0":
            raise ValueError("ERROR (2020): only MODE0 phase-space files are supported")
        if count is None:
            count = npp - start
        if start + count > npp:
            raise ValueError("ERROR (2019): fewer than %d particles remain" % count)
        f.seek(28 + 28 * start)
        rec = np.frombuffer(f.read(28 * count), dtype=PHSP_DTYPE).copy()
    return dict(npphsp=npp, nphotphsp=nphot, ekmax=ekmax, ekmin=ekmin, ninc=ninc), rec


def decode_phsp(rec, z_phsp):
    """MODE0 records -> source particles (CUDA_EGS.cu:56-106)."""
    out = np.zeros(len(rec), dtype=SOURCE_DTYPE)
    latch = rec["latch"]
    q = np.where(latch & 0x40000000, -1, np.where(latch & 0x20000000, 1, 0))
    u = rec["u"].astype(np.float32)
    v = rec["v"].astype(np.float32)
    w = np.sqrt(np.maximum(np.float32(0.0), np.float32(1.0) - u * u - v * v)).astype(np.float32)
    wt = rec["wt"].astype(np.float32)
    w = np.where(wt < 0, -w, w)
    out["e"] = np.abs(rec["e"])
    out["x"] = rec["x"]
    out["y"] = rec["y"]
    out["z"] = np.float32(z_phsp)
    out["u"] = u
    out["v"] = v
    out["w"] = w
    out["wt"] = np.abs(wt)
    out["charge"] = q
    return out


def parallel_beam(n, energy, half_width, z0, seed):
    """Parallel square field of monoenergetic photons travelling along +z."""
    rng = np.random.default_rng(seed)
    s = np.zeros(n, dtype=SOURCE_DTYPE)
    s["e"] = np.float32(energy)
    s["x"] = rng.uniform(-half_width, half_width, n).astype(np.float32)
    s["y"] = rng.uniform(-half_width, half_width, n).astype(np.float32)
    s["z"] = np.float32(z0)
    s["w"] = np.float32(1.0)
    s["wt"] = np.float32(1.0)
    return s


def mv6_spectrum(nbins=24, emin=0.25, emax=6.0):
    """Synthetic 6 MV-like bremsstrahlung spectrum: a thin-target (Kramers)
    shape k^-1 (E0 - k) filtered by 1.5 cm of tungsten-equivalent attenuation,
    normalised; 24 bins over 0.25-6 MeV.  Returns (bin edges, probabilities)."""
    edges = np.linspace(emin, emax, nbins + 1)
    mid = 0.5 * (edges[1:] + edges[:-1])
    thin = (emax + 0.2 - mid) / mid
    mu = 0.04 + 0.9 * mid ** -1.9 * 0.02      # crude flattening-filter hardening
    w = thin * np.exp(-mu * 8.0)
    w /= w.sum()
    return edges, w


def point_source_spectrum(n, ssd, field, seed, z_source=0.0, edges=None, probs=None):
    """Point source at the origin of the head frame; directions towards points
    uniform over a field x field square at distance ``ssd``; energies from the
    binned spectrum (uniform within a bin)."""
    if edges is None:
        edges, probs = mv6_spectrum()
    rng = np.random.default_rng(seed)
    half = field / 2.0
    px = rng.uniform(-half, half, n)
    py = rng.uniform(-half, half, n)
    norm = np.sqrt(px * px + py * py + ssd * ssd)
    s = np.zeros(n, dtype=SOURCE_DTYPE)
    b = rng.choice(len(probs), size=n, p=probs)
    e = edges[b] + rng.uniform(0.0, 1.0, n) * (edges[b + 1] - edges[b])
    s["e"] = e.astype(np.float32)
    s["x"] = np.float32(0.0)
    s["y"] = np.float32(0.0)
    s["z"] = np.float32(z_source)
    s["u"] = (px / norm).astype(np.float32)
    s["v"] = (py / norm).astype(np.float32)
    s["w"] = (ssd / norm).astype(np.float32)
    s["wt"] = np.float32(1.0)
    return s


def encode_phsp(src):
    """Source particles -> MODE0 records (the inverse of decode_phsp): charge
    in latch bits 29/30, total energy (EGSnrc stores it for charged particles,
    the photon energy otherwise), w < 0 flagged by a negative weight."""
    rec = np.zeros(len(src), dtype=PHSP_DTYPE)
    q = src["charge"]
    rec["latch"] = np.where(q < 0, 0x40000000, np.where(q > 0, 0x20000000, 0)).astype(np.uint32)
    rec["e"] = src["e"]
    rec["x"] = src["x"]
    rec["y"] = src["y"]
    rec["u"] = src["u"]
    rec["v"] = src["v"]
    rec["wt"] = np.where(src["w"] < 0, -src["wt"], src["wt"]).astype(np.float32)
    return rec


def head_phase_space(n, z_plane, ssd, field, seed, electron_fraction=0.03, edges=None, probs=None, center=(0.0, 0.0)):
    """Synthetic phase space scored on the plane z = z_plane above the
    collimators: particles from a point source at the origin travelling
    towards points uniform over a (field + 2 cm) square at distance ssd;
    photons with the 6 MV-like spectrum, a fraction of contaminant electrons
    (kinetic energy uniform in 0.2-4 MeV, total energy stored).  The square
    is centred on ``center`` (x, y at distance ssd)."""
    if edges is None:
        edges, probs = mv6_spectrum()
    rng = np.random.default_rng(seed)
    half = field / 2.0 + 1.0
    px = rng.uniform(-half, half, n) + center[0]
    py = rng.uniform(-half, half, n) + center[1]
    norm = np.sqrt(px * px + py * py + ssd * ssd)
    u, v, w = px / norm, py / norm, ssd / norm
    t = z_plane / w
    s = np.zeros(n, dtype=SOURCE_DTYPE)
    b = rng.choice(len(probs), size=n, p=probs)
    e = edges[b] + rng.uniform(0.0, 1.0, n) * (edges[b + 1] - edges[b])
    el = rng.uniform(0.0, 1.0, n) < electron_fraction
    s["e"] = np.where(el, rng.uniform(0.2, 4.0, n) + RM, e).astype(np.float32)
    s["charge"] = np.where(el, -1, 0)
    s["x"] = (u * t).astype(np.float32)
    s["y"] = (v * t).astype(np.float32)
    s["z"] = np.float32(z_plane)
    s["u"] = u.astype(np.float32)
    s["v"] = v.astype(np.float32)
    s["w"] = w.astype(np.float32)
    s["wt"] = np.float32(1.0)
    return s


def stratified_parallel_beam(n_side, energy, half_width, z0, seed):
    """Parallel square field of monoenergetic photons along +z with one source
    point in each cell of an n_side x n_side grid over the field (uniform
    within its cell): the source positions add no sampling noise beyond the
    cell size, so reusing the records for many histories (different history
    offsets) leaves only the transport's own noise."""
    rng = np.random.default_rng(seed)
    n = n_side * n_side
    cell = 2.0 * half_width / n_side
    idx = np.arange(n, dtype=np.int64)
    s = np.zeros(n, dtype=SOURCE_DTYPE)
    s["e"] = np.float32(energy)
    s["x"] = (-half_width + cell * ((idx % n_side) + rng.random(n))).astype(np.float32)
    s["y"] = (-half_width + cell * ((idx // n_side) + rng.random(n))).astype(np.float32)
    s["z"] = np.float32(z0)
    s["w"] = np.float32(1.0)
    s["wt"] = np.float32(1.0)
    return s


# ------------------------------------------------------------------ the run input's phase-space files
class PhaseSpaceFiles:
    """The phase-space input of a run: ``n_files`` MODE0 files
    ``<base>_p<i>.egsphsp1`` (i = 0 .. n_files - 1), or one file, read as one
    record sequence (read_egsphsp_header / read_egsphsp_data,
    read_egsphsp.cpp:60-232).

    Kept from the reference (each with a switch that corrects it):

    * Q4: with one file the header (its record count) comes from
      ``<base>.egsphsp1`` but the records from ``<base>_p0.egsphsp1``
      (read_egsphsp.cpp:113 against :202-216); ``fix_q4`` reads both from
      ``<base>.egsphsp1``.
    * Q15: a read that runs past the end of file i takes its remaining
      records from the start of file i again, not of file i + 1
      (read_egsphsp.cpp:208-212 open the same ``_p<i>`` name twice);
      ``fix_q15`` continues in file i + 1.

    ``read(istart, length)`` returns None where the reference returns -1:
    ``istart`` at or past the total, or fewer than ``length`` records left."""

    def __init__(self, base, n_files, fix_q4=False, fix_q15=False):
        self.base = base
        self.n_files = int(n_files)
        self.fix_q15 = fix_q15
        if self.n_files > 1:
            self.header_paths = ["%s_p%d.egsphsp1" % (base, i) for i in range(self.n_files)]
            self.data_paths = list(self.header_paths)
        else:
            self.header_paths = [base + ".egsphsp1"]
            self.data_paths = [base + ".egsphsp1" if fix_q4 else base + "_p0.egsphsp1"]
        self.counts = []
        for p in self.header_paths:
            try:
                with open(p, "rb") as f:
                    hdr = f.read(28)
            except OSError:
                raise ValueError('Could not open the egsphsp file "%s".' % p)
            if len(hdr) < 28:
                raise ValueError('The egsphsp file "%s" has no complete header.' % p)
            self.counts.append(struct.unpack("<5s3xI", hdr[:12])[1])
        self.total = int(sum(self.counts))

    def _records(self, i, start, count):
        path = self.data_paths[i]
        if not os.path.exists(path):
            raise ValueError('Could not open the egsphsp file "%s"%s.' % (
                path, " (the reference reads the records of a single phase-space file from <name>_p0.egsphsp1,"
                      " quirk Q4; fix_q4 reads <name>.egsphsp1)" if self.n_files <= 1 else ""))
        with open(path, "rb") as f:
            f.seek(28 + 28 * start)
            raw = f.read(28 * count)
        if len(raw) != 28 * count:
            raise ValueError('The egsphsp file "%s" holds fewer than %d records from record %d.' % (path, count, start))
        return np.frombuffer(raw, dtype=PHSP_DTYPE).copy()

    def read(self, istart, length):
        istart, length = int(istart), int(length)
        if istart >= self.total or self.total - istart < length:
            return None
        i = 0
        while i < len(self.counts) and istart >= self.counts[i]:
            istart -= self.counts[i]
            i += 1
        if i >= len(self.counts):
            return None
        remain = self.counts[i] - istart
        if remain < length:
            second = i + 1 if self.fix_q15 else i
            return np.concatenate([self._records(i, istart, remain), self._records(second, 0, length - remain)])
        return self._records(i, istart, length)
