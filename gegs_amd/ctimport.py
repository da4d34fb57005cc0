"""CT front end: a DICOM CT series to a transport phantom.

Restates the reference's chain (ctcreate.c:1342-1600): ``readct_dicom``
(ReadCT_DICOM.c:316-909) reads the slices named in a list file, orders them
by the z of their image position, drops localizers and slices at a z already
seen, takes the slice thickness from the first two slices and applies the
first slice's rescale slope and intercept; ``SubsizeCT`` crops to a box;
``ResampleCT`` averages the CT numbers onto the transport voxels;
``CTToMedium`` maps them to media and densities through the ramp.  Cropping
and resampling run in the library (gegs_ct_subsize, gegs_ct_resample) in the
reference's binary32 arithmetic and order; the ramp is ``phantom.ct_to_medium``.

The DICOM reader parses what the reference parses: little-endian files with
or without the 128-byte preamble and "DICM", explicit or implicit VR decided
element by element by whether the two bytes after the tag name a VR, 16-bit
pixel data.  Lengths are in cm.  Layout: x runs along a row (the column
index), y along the columns (the row index), z over slices; the x spacing is
the column spacing.  For square images with square pixels -- the only case
the reference's fixed-stride arrays handle (CTIMAX = CTJMAX = 512,
EGSnrc_Parameters.h:137-138) -- this is the reference's layout.
"""
from __future__ import annotations

import ctypes
import io
import math
import os
import struct

import numpy as np

from . import phantom as ph

VRS = {b"AE", b"AS", b"AT", b"CS", b"DA", b"DS", b"DT", b"FL", b"FD", b"IS", b"LO", b"LT", b"OB", b"OW",
       b"PN", b"SH", b"SL", b"SQ", b"SS", b"ST", b"TM", b"UI", b"UL", b"UN", b"US", b"UT"}
LONG_VRS = {b"OB", b"OW", b"SQ", b"UN"}

TAG_IMAGE_TYPE = 0x00080008
TAG_IMAGE_POSITION = 0x00200032
TAG_IMAGE_ORIENTATION = 0x00200037
TAG_ROWS = 0x00280010
TAG_COLUMNS = 0x00280011
TAG_PIXEL_SPACING = 0x00280030
TAG_RESCALE_INTERCEPT = 0x00281052
TAG_RESCALE_SLOPE = 0x00281053
TAG_PIXEL_DATA = 0x7FE00010


def elements(data):
    """(tag, value bytes) of every element with a defined, non-zero length,
    in file order (read_elem_vr / read_val, ReadCT_DICOM.c:138-280).
    Elements of undefined length (sequences) are entered, not skipped: their
    items are read as further elements.  The 128-byte preamble is consumed as
    sixteen zero elements followed by "DICM"."""
    pos, n, zeros = 0, len(data), 0
    while pos + 4 <= n:
        group, elem = struct.unpack_from("<HH", data, pos)
        tag = (group << 16) | elem
        pos += 4
        if pos + 2 > n:
            return
        vr = bytes(data[pos:pos + 2])
        if vr in VRS:
            if vr in LONG_VRS:
                if pos + 8 > n:
                    return
                length = struct.unpack_from("<I", data, pos + 4)[0]
                pos += 8
            else:
                if pos + 4 > n:
                    return
                length = struct.unpack_from("<H", data, pos + 2)[0]
                pos += 4
        else:
            if pos + 4 > n:
                return
            length = struct.unpack_from("<I", data, pos)[0]
            pos += 4
        if length == 0 and tag == 0:
            zeros += 1
        if zeros == 16:
            pos += 4                          # "DICM"
            zeros = 0
        elif 0 < length < 0xFFFFFFFF:
            if pos + length > n:
                return
            yield tag, bytes(data[pos:pos + length])
            pos += length


def _bad(what):
    return ValueError("ERROR (5008): DICOM series: %s" % what)


def _decimal(text):
    """One finite decimal (surrounding blanks allowed), else error 5008."""
    try:
        v = float(text)
    except ValueError:
        raise _bad("%r is not a number" % text) from None
    if not math.isfinite(v):
        raise _bad("%r is not finite" % text)
    return v


def _doubles(value, count):
    """Backslash-separated decimal strings (string_to_doubles)."""
    parts = value.rstrip(b"\x00 ").split(b"\\")
    if len(parts) < count:
        raise _bad("expected %d values in %r" % (count, value))
    return [_decimal(p) for p in parts[:count]]


def _read(path, mode="rb"):
    try:
        with open(path, mode) as f:
            return f.read()
    except OSError:
        raise _bad("cannot read %s" % path) from None


def _u16(val):
    if len(val) < 2:
        raise _bad("a rows or columns element shorter than 2 bytes")
    return struct.unpack_from("<H", val)[0]


class CTVolume:
    """CT numbers hu[k, j, i] (int16, x fastest), voxel sizes and the lower
    corner (cm)."""

    def __init__(self, hu, voxel, offset):
        self.hu = np.ascontiguousarray(hu, dtype=np.int16)
        self.voxel = np.asarray(voxel, dtype=np.float32)
        self.offset = np.asarray(offset, dtype=np.float32)

    @property
    def size(self):
        nz, ny, nx = self.hu.shape
        return np.array([nx, ny, nz], dtype=np.int32)


def read_dicom_series(list_path):
    """Read the slices named (one per line, relative to the list's folder)
    in list_path; returns a CTVolume.  Follows readct_dicom: slices ordered by
    increasing z of their image position, localizers and repeated z dropped,
    dz = z[1] - z[0], rows/columns/spacing/rescale from the first slice,
    rescale applied only when slope and intercept are both non-zero, lower
    corner = first pixel centre - half a voxel."""
    folder = os.path.dirname(os.path.abspath(list_path))
    names = [line.rstrip("\r\n") for line in io.StringIO(_read(list_path, "r"))]
    names = [nm for nm in names if nm.strip()]
    slices = []                     # (z, order, path)
    seen = set()
    for order, nm in enumerate(names):
        path = os.path.join(folder, nm)
        data = _read(path)
        z = None
        for tag, val in elements(data):
            if tag == TAG_IMAGE_TYPE and b"LOCALIZER" in val:
                break
            if tag == TAG_IMAGE_POSITION:
                z = np.float32(_doubles(val, 3)[2] / 10.0)
                break
        if z is None:
            continue
        if float(z) in seen:
            continue                # a slice at a z already read
        seen.add(float(z))
        slices.append((float(z), order, path))
    if len(slices) < 2:
        raise _bad("need at least two CT slices at distinct z")
    slices.sort()
    dz = np.float32(np.float32(slices[1][0]) - np.float32(slices[0][0]))
    rows = cols = None
    offset = np.zeros(3, dtype=np.float32)
    spacing = None
    slope = intercept = np.float32(0.0)
    planes = []
    for n, (_, _, path) in enumerate(slices):
        data = _read(path)
        pixels = None
        for tag, val in elements(data):
            if n == 0:
                if tag == TAG_ROWS:
                    rows = _u16(val)
                elif tag == TAG_COLUMNS:
                    cols = _u16(val)
                elif tag == TAG_IMAGE_POSITION:
                    p = _doubles(val, 3)
                    offset = np.array([p[0] / 10.0, p[1] / 10.0, p[2] / 10.0], dtype=np.float32)
                elif tag == TAG_PIXEL_SPACING:
                    spacing = _doubles(val, 2)
                elif tag == TAG_RESCALE_INTERCEPT:
                    intercept = np.float32(_decimal(val.rstrip(b"\x00 ")))
                elif tag == TAG_RESCALE_SLOPE:
                    slope = np.float32(_decimal(val.rstrip(b"\x00 ")))
            if tag == TAG_PIXEL_DATA:
                if len(val) % 2:
                    raise _bad("%s: pixel data of odd length" % path)
                pixels = np.frombuffer(val, dtype="<i2")
        if rows is None or cols is None or spacing is None or pixels is None:
            raise _bad("%s: missing rows, columns, pixel spacing or pixel data" % path)
        if rows == 0 or cols == 0:
            raise _bad("%s: no rows or no columns" % path)
        if pixels.size < rows * cols:
            raise _bad("%s: short pixel data" % path)
        planes.append(pixels[:rows * cols].reshape(rows, cols))
    hu = np.stack(planes).astype(np.int16)
    if slope != 0 and intercept != 0:
        # binary32 product and sum, stored to a 16-bit integer through a
        # truncating conversion to 32 bits, as the reference's float to short
        # assignment (ReadCT_DICOM.c:879-882) compiles on x86: values past
        # 32 bits become INT_MIN, the 16-bit store keeps the low half
        v = (hu.astype(np.float32) * slope + intercept).astype(np.float64)
        t = np.trunc(v)
        ok = np.isfinite(t) & (t >= -2.0 ** 31) & (t < 2.0 ** 31)
        hu = np.where(ok, t, -2.0 ** 31).astype(np.int64).astype(np.int32).astype(np.int16)
    voxel = np.array([spacing[1] / 10.0, spacing[0] / 10.0, 0.0], dtype=np.float32)
    voxel[2] = dz
    corner = np.array([np.float64(offset[a]) - np.float64(voxel[a]) / 2.0 for a in range(3)], dtype=np.float32)
    return CTVolume(hu, voxel, corner)


def _lib():
    from . import native
    return native.load()


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def _i3(v):
    return (ctypes.c_int32 * 3)(*[int(x) for x in v])


def subsize(vol, box_min=(0, 0, 0), box_max=(0, 0, 0)):
    """SubsizeCT: the CT voxels covering the box (cm; all zero = whole CT)."""
    L = _lib()
    hu = np.ascontiguousarray(vol.hu.copy())
    size = _i3(vol.size)
    off = _f3(vol.offset)
    rc = L.gegs_ct_subsize(size, ctypes.c_void_p(hu.ctypes.data), _f3(vol.voxel), off, _f3(box_min), _f3(box_max))
    if rc:
        raise ValueError("gegs_ct_subsize: error %d" % rc)
    nx, ny, nz = size[0], size[1], size[2]
    out = hu.reshape(-1)[:nx * ny * nz].reshape(nz, ny, nx).copy()
    return CTVolume(out, vol.voxel, [off[0], off[1], off[2]])


def resample(vol, thickness):
    """ResampleCT: (hu[k, j, i], xb, yb, zb) on voxels of about `thickness` cm."""
    L = _lib()
    t = _f3(thickness)
    out = (ctypes.c_int32 * 3)()
    size = _i3(vol.size)
    rc = L.gegs_ct_resample_grid(size, _f3(vol.voxel), t, out)
    if rc:
        raise ValueError("ResampleCT: error %d" % rc)
    nx, ny, nz = out[0], out[1], out[2]
    xb = np.zeros(nx + 1, dtype=np.float32)
    yb = np.zeros(ny + 1, dtype=np.float32)
    zb = np.zeros(nz + 1, dtype=np.float32)
    res = np.zeros((nz, ny, nx), dtype=np.int16)
    hu = np.ascontiguousarray(vol.hu)
    rc = L.gegs_ct_resample(size, _f3(vol.voxel), ctypes.c_void_p(hu.ctypes.data), _f3(vol.offset), t, out,
                            ctypes.c_void_p(xb.ctypes.data), ctypes.c_void_p(yb.ctypes.data),
                            ctypes.c_void_p(zb.ctypes.data), ctypes.c_void_p(res.ctypes.data))
    if rc:
        raise ValueError("gegs_ct_resample: error %d" % rc)
    return res, xb, yb, zb


def read_ct_ramps(path, max_media=6):
    """CT ramps file (ctcreate.c:886-1040): "n,lowest CT number" on the first
    line, then per medium a name line (24 characters) and "CT upper bound,
    density at the lower end, density at the upper end, unused".  n = 0
    selects the built-in 4-media ramp.  Returns (ramp, lowest CT number);
    error 5005 for n outside [0, max_media] (the reference's MXMED is 6,
    EGSnrc_Parameters.h:5), 5006 for a medium whose CT upper bound is below
    the lowest CT number, whose lower density is negative or whose upper
    density is below its lower one (ctcreate.c:1036-1043), 5007 for a file
    that cannot be read or parsed (the reference reads it unchecked)."""
    def bad(what):
        return ValueError("ERROR (5007): CT ramps file %s: %s" % (path, what))

    try:
        with open(path) as f:
            lines = f.read().splitlines()
    except OSError:
        raise bad("cannot be read") from None
    try:
        head = lines[0].replace(" ", "").split(",")
        n, lower = int(head[0]), int(head[1])
    except (IndexError, ValueError):
        raise bad("no \"number of media, lowest CT number\" line") from None
    if n < 0 or n > max_media:
        raise ValueError("ERROR (5005): number of materials out of range")
    if n == 0:
        return list(ph.DEFAULT_RAMP), ph.DEFAULT_RAMP_LOWER
    ramp = []
    for m in range(n):
        try:
            name = lines[1 + 2 * m][:24].strip()
            vals = [v for v in lines[2 + 2 * m].replace(" ", "").split(",") if v]
            up, dlo, dhi = int(vals[0]), float(np.float32(float(vals[1]))), float(np.float32(float(vals[2])))
        except (IndexError, ValueError):
            raise bad("medium %d needs a name line and a \"CT upper bound, density low, density high\" line"
                      % (m + 1)) from None
        if up < lower or dlo < 0 or dhi < dlo:
            raise ValueError("ERROR (5006): CT no. or density out of range, or density upper bound < "
                             "density lower bound (medium %d)" % (m + 1))
        ramp.append((name, up, dlo, dhi))
    return ramp, lower


def ctcreate(list_path, thickness, box_min=(0, 0, 0), box_max=(0, 0, 0), ramp=ph.DEFAULT_RAMP,
             lower=ph.DEFAULT_RAMP_LOWER, fix_q10=False):
    """DICOM series -> Phantom: read, crop, resample, CT ramp (fix_q10: CT
    values below the ramp take its first medium instead of vacuum)."""
    vol = subsize(read_dicom_series(list_path), box_min, box_max)
    hu, xb, yb, zb = resample(vol, thickness)
    med, dens = ph.ct_to_medium(hu.reshape(-1), ramp=ramp, lower=lower, fix_q10=fix_q10)
    return ph.Phantom([r[0] for r in ramp], xb, yb, zb, med, dens)


def write_egsphant_ctcreate(path, ph):
    """The .egsphant file ctcreate writes (ctcreate.c:1249-1311): the media
    count "%2d", the names, a dummy ESTEPE line " 1.0" per medium, " %4d" x3
    voxel counts, each axis's bounds as " %g" eight to a line, the media
    digits a row per line with an empty line after each slice, then the
    densities as "  %g" eight to a line (counted over a slice) with an empty
    line after each slice.  The reference reads the phantom back from this
    text, so its bounds and densities keep the 6 significant digits of %g."""
    nx, ny, nz = ph.nx, ph.ny, ph.nz
    out = ["%2d\n" % len(ph.media_names)]
    out += ["%s\n" % n for n in ph.media_names]
    out.append("".join(" %3.1f" % 1.0 for _ in ph.media_names) + "\n")
    out.append(" %4d %4d %4d\n" % (nx, ny, nz))
    for b, n in ((ph.xb, nx), (ph.yb, ny), (ph.zb, nz)):
        for ii in range(n + 1):
            out.append(" %g" % float(b[ii]))
            if ii % 8 == 7:
                out.append("\n")
        if n % 8 != 7:
            out.append("\n")
    med = np.asarray(ph.media_idx).reshape(nz, ny, nx)
    for k in range(nz):
        for j in range(ny):
            out.append("".join("%1d" % v for v in med[k, j]) + "\n")
        out.append("\n")
    rho = np.asarray(ph.density, dtype=np.float32).reshape(nz, ny, nx)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                out.append("  %g" % float(rho[k, j, i]))
                if (i + j * nx) % 8 == 7:
                    out.append("\n")
        if (nx * ny) % 8 != 0:
            out.append("\n")
        out.append("\n")
    with open(path, "w") as f:
        f.write("".join(out))


def phantom_from_run_input(ri, base_dir=None):
    """The phantom a run input asks for with "use ctcreate = Yes" ([CTcreate]:
    CT ramps file, dicom file list, xyz boundaries x1 x2 y1 y2 z1 z2, 3d voxel
    sizes), files relative to base_dir (default: the input's folder)."""
    if not ri.use_ctcreate:
        raise ValueError("the run input does not use ctcreate")
    base = base_dir or os.path.dirname(os.path.abspath(ri.path))

    def local(p):
        p = p.replace("\\", "/")
        return p if os.path.isabs(p) else os.path.join(base, p)

    ramp, lower = read_ct_ramps(local(ri.ct_ramps_file))
    b = ri.xyz_boundaries
    return ctcreate(local(ri.dicom_file_list), ri.ct_voxel_sizes, (b[0], b[2], b[4]), (b[1], b[3], b[5]),
                    ramp=ramp, lower=lower)


# ------------------------------------------------------------------ writer
def _element(group, elem, vr, value):
    if vr in LONG_VRS:
        return struct.pack("<HH2sHI", group, elem, vr, 0, len(value)) + value
    return struct.pack("<HH2sH", group, elem, vr, len(value)) + value


def _text(s):
    b = s.encode("ascii")
    return b + (b" " if len(b) % 2 else b"")


def write_ct_slice(path, pixels, position_mm, spacing_mm, slope=1.0, intercept=-1024.0, localizer=False,
                   implicit=False, preamble=True):
    """One CT slice as a little-endian DICOM file (explicit VR, or implicit VR
    for the data set), with the elements the reader uses."""
    rows, cols = pixels.shape
    img_type = "ORIGINAL\\PRIMARY\\LOCALIZER" if localizer else "ORIGINAL\\PRIMARY\\AXIAL"
    meta = _element(0x0002, 0x0010, b"UI", _text("1.2.840.10008.1.2" if implicit else "1.2.840.10008.1.2.1"))
    ds = [
        (0x0008, 0x0008, b"CS", _text(img_type)),
        (0x0020, 0x0032, b"DS", _text("%.6f\\%.6f\\%.6f" % tuple(position_mm))),
        (0x0020, 0x0037, b"DS", _text("1\\0\\0\\0\\1\\0")),
        (0x0028, 0x0010, b"US", struct.pack("<H", rows)),
        (0x0028, 0x0011, b"US", struct.pack("<H", cols)),
        (0x0028, 0x0030, b"DS", _text("%.6f\\%.6f" % tuple(spacing_mm))),
        (0x0028, 0x1052, b"DS", _text("%g" % intercept)),
        (0x0028, 0x1053, b"DS", _text("%g" % slope)),
        (0x7FE0, 0x0010, b"OW", np.ascontiguousarray(pixels, dtype="<i2").tobytes()),
    ]
    body = b""
    for g, e, vr, val in ds:
        if implicit:
            body += struct.pack("<HHI", g, e, len(val)) + val
        else:
            body += _element(g, e, vr, val)
    with open(path, "wb") as f:
        if preamble:
            f.write(b"\x00" * 128 + b"DICM")
        f.write(meta + body)
